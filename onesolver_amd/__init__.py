"""Python binding of libonesolver.so, the MI355X simulated-annealing QUBO library.

Thin ctypes layer over the C ABI declared in ``include/onesolver.h``; used by the
tests and the benchmark.  It never falls back to a CPU implementation: if the
library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libonesolver.so")
CLI_PATH = os.path.join(HERE, "bin", "one-solver-anneal")
BATCH_CLI_PATH = os.path.join(HERE, "bin", "one-solver-batch")
EXHAUSTIVE_CLI_PATH = os.path.join(HERE, "bin", "one-solver-exhaustive")
TEST_BIN_PATH = os.path.join(HERE, "bin", "main-test")

OS_OK = 0
OS_ERR_INVALID = -1
OS_ERR_HIP = -2
OS_ERR_UNSUPPORTED = -3

_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p

_lib = None


class OneSolverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.message = msg


class InvalidArgument(OneSolverError, ValueError):
    pass


def lib():
    """Load libonesolver.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libonesolver.so not built: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        L.os_version.restype = ctypes.c_int
        L.os_last_error.restype = ctypes.c_char_p
        L.os_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.os_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.os_ctx_destroy.argtypes = [_vp]
        L.os_ctx_device_name.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t]
        L.os_anneal_reference.argtypes = [_vp, ctypes.c_int, _f64p, _f64p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, _u8p,
                                          _f64p, _u8p, _f64p]
        L.os_energy_fp64.argtypes = [_vp, ctypes.c_int, _f64p, ctypes.c_uint32, _u8p, _f64p]
        L.os_exhaustive.argtypes = [_vp, ctypes.c_int, _f64p, _u8p, _f64p]
        L.os_dense_create.argtypes = [_vp, ctypes.c_int, _f64p, ctypes.POINTER(_vp)]
        L.os_dense_destroy.argtypes = [_vp]
        L.os_dense_anneal.argtypes = [_vp, ctypes.POINTER(SweepParams), _f64p, _u8p, _u8p, _u8p,
                                      _f64p, ctypes.POINTER(SweepStats)]
        _lib = L
    return _lib


class SweepParams(ctypes.Structure):
    """os_sweep_params of include/onesolver.h."""
    _fields_ = [("num_replicas", ctypes.c_uint32), ("num_sweeps", ctypes.c_int32),
                ("tscale", _vp), ("seed", ctypes.c_uint64), ("instance", ctypes.c_uint32),
                ("refresh_interval", ctypes.c_int32), ("init_states", _vp),
                ("init_marginals", _vp)]


class TemperingParams(ctypes.Structure):
    """os_tempering_params of include/onesolver.h."""
    _fields_ = [("num_ladders", ctypes.c_uint32), ("num_temperatures", ctypes.c_uint32),
                ("tscale", _vp), ("num_sweeps", ctypes.c_int32),
                ("exchange_interval", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("instance", ctypes.c_uint32), ("refresh_interval", ctypes.c_int32)]


class PopulationParams(ctypes.Structure):
    """os_population_params of include/onesolver.h."""
    _fields_ = [("num_replicas", ctypes.c_uint32), ("num_steps", ctypes.c_int32),
                ("sweeps_per_step", ctypes.c_int32), ("tscale", _vp), ("seed", ctypes.c_uint64),
                ("instance", ctypes.c_uint32), ("refresh_interval", ctypes.c_int32)]


class SweepStats(ctypes.Structure):
    """os_sweep_stats of include/onesolver.h."""
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("kernel_launches", ctypes.c_int32), ("max_field_drift", ctypes.c_double),
                ("attempts", ctypes.c_uint64), ("pass_avg_ms", ctypes.c_double),
                ("pass_launches", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def tscale_from_beta(beta, temperature=False):
    """Per-sweep threshold scale: 1/beta (inverse temperature) or beta (the
    reference's use of beta as a temperature)."""
    b = np.asarray(beta, dtype=np.float64)
    return (b if temperature else 1.0 / b).astype(np.float32)


def _check(rc):
    if rc != OS_OK:
        msg = lib().os_last_error().decode(errors="replace")
        if rc == OS_ERR_INVALID:
            raise InvalidArgument(rc, msg)
        raise OneSolverError(rc, msg)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(lib().os_device_count(ctypes.byref(c)))
    return c.value


class Context:
    """One GPU, one HIP stream (``os_ctx``)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().os_ctx_create(device, ctypes.byref(h)))
        self.device = device
        self._h = h
        # objects created on this context: closed before it is destroyed, so
        # none of them outlives the os_ctx it refers to
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        if not self._h:
            raise OneSolverError(-1, "context is closed")
        return self._h

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self._h:
            for child in list(self._children):
                child.close()
            lib().os_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(lib().os_ctx_device_name(self._h, buf, 256))
        return buf.value.decode()

    def anneal_reference(self, q_sym, beta, num_tries, sweeps_per_beta=1, seed=1234,
                         want_states=True):
        """sa::anneal semantics; returns (best_state, best_energy, states, energies)."""
        q = np.ascontiguousarray(q_sym, dtype=np.float64)
        n = q.shape[0]
        b = np.ascontiguousarray(beta, dtype=np.float64)
        states = np.zeros((num_tries, n), dtype=np.uint8) if want_states else None
        energies = np.zeros(num_tries, dtype=np.float64)
        best = np.zeros(n, dtype=np.uint8)
        be = ctypes.c_double(0.0)
        _check(lib().os_anneal_reference(self._h, n, _p(q, _f64p), _p(b, _f64p), len(b),
                                         sweeps_per_beta, num_tries, seed, _p(states, _u8p),
                                         _p(energies, _f64p), _p(best, _u8p), ctypes.byref(be)))
        return best, be.value, states, energies

    def energy_fp64(self, q_sym, states):
        q = np.ascontiguousarray(q_sym, dtype=np.float64)
        n = q.shape[0]
        st = np.ascontiguousarray(states, dtype=np.uint8).reshape(-1, n)
        out = np.zeros(st.shape[0], dtype=np.float64)
        _check(lib().os_energy_fp64(self._h, n, _p(q, _f64p), st.shape[0], _p(st, _u8p),
                                    _p(out, _f64p)))
        return out


    def exhaustive(self, q_upper):
        """(state, energy) of the minimum over all 2^n states of the upper
        triangle of q_upper, lowest state index on ties (os_exhaustive)."""
        q = np.ascontiguousarray(q_upper, dtype=np.float64)
        n = q.shape[0]
        st = np.zeros(n, dtype=np.uint8)
        e = ctypes.c_double(0.0)
        _check(lib().os_exhaustive(self._h, n, _p(q, _f64p), _p(st, _u8p), ctypes.byref(e)))
        return st, e.value


def _dense_tune(h, num_replicas, tscale, seed, instance):
    """os_dense_tune: the streaming path's blocks per field pass and pass
    work-groups per CU, timed on probe runs (zeros for n <= 1024)."""
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    prm = SweepParams(int(num_replicas), len(ts), ts.ctypes.data, seed, instance, 0, None, None)
    g, w, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0.0)
    _check(lib().os_dense_tune(h, ctypes.byref(prm), ctypes.byref(g), ctypes.byref(w),
                               ctypes.byref(ms)))
    return {"group": g.value, "pass_wgs_per_cu": w.value, "probe_ms": ms.value}


class DenseProblem:
    """A dense QUBO resident on the GPU for sweep-mode annealing (os_dense)."""

    def __init__(self, ctx: Context, q_sym):
        q = np.ascontiguousarray(q_sym, dtype=np.float64)
        self.n = q.shape[0]
        self.ctx = ctx
        h = _vp()
        _check(lib().os_dense_create(ctx.handle, self.n, _p(q, _f64p), ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    def close(self):
        if self._h:
            lib().os_dense_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def anneal(self, num_replicas, tscale, seed=1, instance=0, refresh_interval=0,
               init_states=None, init_marginals=None, best_states=True, final_states=False):
        ts = np.ascontiguousarray(tscale, dtype=np.float32)
        n = self.n
        ini = None if init_states is None else np.ascontiguousarray(init_states, dtype=np.uint8)
        mar = None if init_marginals is None else np.ascontiguousarray(init_marginals, dtype=np.float32)
        prm = SweepParams(num_replicas, len(ts), ts.ctypes.data, seed, instance, refresh_interval,
                          None if ini is None else ini.ctypes.data,
                          None if mar is None else mar.ctypes.data)
        be = np.zeros(num_replicas, dtype=np.float64)
        bs = np.zeros((num_replicas, n), dtype=np.uint8) if best_states else None
        fs = np.zeros((num_replicas, n), dtype=np.uint8) if final_states else None
        best = np.zeros(n, dtype=np.uint8)
        e = ctypes.c_double(0.0)
        st = SweepStats()
        _check(lib().os_dense_anneal(self._h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p),
                                     _p(fs, _u8p), _p(best, _u8p), ctypes.byref(e),
                                     ctypes.byref(st)))
        return {"best_energies": be, "best_states": bs, "final_states": fs, "best_state": best,
                "best_energy": e.value, "stats": st.as_dict()}

    def tune(self, num_replicas, tscale, seed=1, instance=0):
        """Autotune the streaming path for these replicas (os_dense_tune)."""
        return _dense_tune(self._h, num_replicas, tscale, seed, instance)

    def tempering(self, ladders, tscale, num_sweeps, exchange_interval, seed=1, instance=0,
                  refresh_interval=0, best_states=True, final_states=False):
        """Replica exchange (os_dense_tempering): `ladders` ladders of
        len(tscale) temperature slots; arrays indexed by replica."""
        out = _tempering(self._h, 1, self.n, ladders, tscale, num_sweeps, exchange_interval, seed,
                         instance, refresh_interval, best_states, final_states)
        for k in ("best_energies", "best_states", "final_states", "final_slots", "best_state"):
            if out[k] is not None:
                out[k] = out[k][0]
        out["best_energy"] = float(out["best_energy"][0])
        return out

    def population(self, num_replicas, tscale, sweeps_per_step, seed=1, instance=0,
                   refresh_interval=0, best_states=True, final_states=False):
        """Population annealing (os_dense_population): num_replicas replicas
        through the len(tscale) temperature steps; arrays indexed by replica,
        ancestors [steps - 1, R]."""
        out = _population(self._h, 1, self.n, num_replicas, tscale, sweeps_per_step, seed,
                          instance, refresh_interval, best_states, final_states)
        for k in ("best_energies", "best_states", "final_states", "ancestors", "best_state"):
            if out[k] is not None:
                out[k] = out[k][0]
        out["best_energy"] = float(out["best_energy"][0])
        return out


def _population(h, I, n, num_replicas, tscale, sweeps_per_step, seed, instance, refresh_interval,
                best_states, final_states):
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    R, K = int(num_replicas), len(ts)
    prm = PopulationParams(R, K, sweeps_per_step, ts.ctypes.data, seed, instance,
                           refresh_interval)
    be = np.zeros((I, R), dtype=np.float64)
    bs = np.zeros((I, R, n), dtype=np.uint8) if best_states else None
    fs = np.zeros((I, R, n), dtype=np.uint8) if final_states else None
    anc = np.zeros((I, max(K - 1, 0), R), dtype=np.uint32)
    best = np.zeros((I, n), dtype=np.uint8)
    e = np.zeros(I, dtype=np.float64)
    st = SweepStats()
    L = lib()
    L.os_dense_population.argtypes = [_vp, ctypes.POINTER(PopulationParams), _f64p, _u8p, _u8p,
                                      ctypes.POINTER(ctypes.c_uint32), _u8p, _f64p,
                                      ctypes.POINTER(SweepStats)]
    _check(L.os_dense_population(h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p),
                                 _p(fs, _u8p),
                                 anc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                 _p(best, _u8p), _p(e, _f64p), ctypes.byref(st)))
    return {"best_energies": be, "best_states": bs, "final_states": fs, "ancestors": anc,
            "best_state": best, "best_energy": e, "stats": st.as_dict()}


def _tempering(h, I, n, ladders, tscale, num_sweeps, exchange_interval, seed, instance,
               refresh_interval, best_states, final_states):
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    R = ladders * len(ts)
    prm = TemperingParams(ladders, len(ts), ts.ctypes.data, num_sweeps, exchange_interval, seed,
                          instance, refresh_interval)
    be = np.zeros((I, R), dtype=np.float64)
    bs = np.zeros((I, R, n), dtype=np.uint8) if best_states else None
    fs = np.zeros((I, R, n), dtype=np.uint8) if final_states else None
    sl = np.zeros((I, R), dtype=np.uint32)
    best = np.zeros((I, n), dtype=np.uint8)
    e = np.zeros(I, dtype=np.float64)
    st = SweepStats()
    L = lib()
    L.os_dense_tempering.argtypes = [_vp, ctypes.POINTER(TemperingParams), _f64p, _u8p, _u8p,
                                     ctypes.POINTER(ctypes.c_uint32), _u8p, _f64p,
                                     ctypes.POINTER(SweepStats)]
    _check(L.os_dense_tempering(h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p), _p(fs, _u8p),
                                sl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                _p(best, _u8p), _p(e, _f64p), ctypes.byref(st)))
    return {"best_energies": be, "best_states": bs, "final_states": fs, "final_slots": sl,
            "best_state": best, "best_energy": e, "stats": st.as_dict()}


class DenseBatch:
    """Same-size dense QUBOs annealed together (os_dense_batch_create)."""

    def __init__(self, ctx: Context, q_batch=None, *, sk=None, uniform=None):
        """q_batch: [I, n, n] FP64 matrices; or sk=(I, n, seed, instance_base)
        / uniform=(I, n, seed, instance_base) to generate SK / dense uniform
        instances on the GPU."""
        L = lib()
        L.os_dense_batch_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _f64p,
                                            ctypes.POINTER(_vp)]
        L.os_dense_batch_create_sk.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.os_dense_copy_q.argtypes = [_vp, _f64p]
        L.os_dense_batch_create_uniform.argtypes = L.os_dense_batch_create_sk.argtypes
        self.ctx = ctx
        h = _vp()
        if uniform is not None:
            I, n, seed, base = uniform
            self.I, self.n = int(I), int(n)
            _check(L.os_dense_batch_create_uniform(ctx.handle, self.I, self.n, seed, base,
                                                   ctypes.byref(h)))
        elif sk is not None:
            I, n, seed, base = sk
            self.I, self.n = int(I), int(n)
            _check(L.os_dense_batch_create_sk(ctx.handle, self.I, self.n, seed, base,
                                              ctypes.byref(h)))
        else:
            q = np.ascontiguousarray(q_batch, dtype=np.float64)
            if q.ndim != 3 or q.shape[1] != q.shape[2]:
                raise ValueError("q_batch must be [instances, n, n]")
            self.I, self.n = q.shape[0], q.shape[1]
            _check(L.os_dense_batch_create(ctx.handle, self.I, self.n, _p(q, _f64p),
                                           ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    def q(self):
        """The batch's FP64 matrices [I, n, n] (copied from the GPU)."""
        out = np.zeros((self.I, self.n, self.n), dtype=np.float64)
        _check(lib().os_dense_copy_q(self._h, _p(out, _f64p)))
        return out

    def close(self):
        if self._h:
            lib().os_dense_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tempering(self, ladders, tscale, num_sweeps, exchange_interval, seed=1, instance=0,
                  refresh_interval=0, best_states=True, final_states=False):
        """Replica exchange on every instance (os_dense_tempering); arrays are
        instance-major, [I, R] and [I, R, n]."""
        return _tempering(self._h, self.I, self.n, ladders, tscale, num_sweeps, exchange_interval,
                          seed, instance, refresh_interval, best_states, final_states)

    def population(self, num_replicas, tscale, sweeps_per_step, seed=1, instance=0,
                   refresh_interval=0, best_states=True, final_states=False):
        """Population annealing on every instance (os_dense_population);
        arrays are instance-major, [I, R], [I, R, n] and ancestors
        [I, steps - 1, R]."""
        return _population(self._h, self.I, self.n, num_replicas, tscale, sweeps_per_step, seed,
                           instance, refresh_interval, best_states, final_states)

    def tune(self, num_replicas, tscale, seed=1, instance=0):
        """Autotune the streaming path for these replicas (os_dense_tune)."""
        return _dense_tune(self._h, num_replicas, tscale, seed, instance)

    def anneal(self, num_replicas, tscale, seed=1, instance=0, refresh_interval=0,
               init_states=None, init_marginals=None, best_states=True, final_states=False):
        """Arrays are instance-major: best_energies [I, R], best_states [I, R, n],
        best_state [I, n], best_energy [I]."""
        ts = np.ascontiguousarray(tscale, dtype=np.float32)
        I, n, R = self.I, self.n, int(num_replicas)
        ini = None if init_states is None else np.ascontiguousarray(init_states, dtype=np.uint8)
        mar = None if init_marginals is None else np.ascontiguousarray(init_marginals,
                                                                       dtype=np.float32)
        if ini is not None and ini.size != I * R * n:
            raise ValueError("init_states must be [I, R, n]")
        if mar is not None and mar.size != I * n:
            raise ValueError("init_marginals must be [I, n]")
        prm = SweepParams(R, len(ts), ts.ctypes.data, seed, instance, refresh_interval,
                          None if ini is None else ini.ctypes.data,
                          None if mar is None else mar.ctypes.data)
        be = np.zeros((I, R), dtype=np.float64)
        bs = np.zeros((I, R, n), dtype=np.uint8) if best_states else None
        fs = np.zeros((I, R, n), dtype=np.uint8) if final_states else None
        best = np.zeros((I, n), dtype=np.uint8)
        e = np.zeros(I, dtype=np.float64)
        st = SweepStats()
        _check(lib().os_dense_anneal(self._h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p),
                                     _p(fs, _u8p), _p(best, _u8p), _p(e, _f64p),
                                     ctypes.byref(st)))
        return {"best_energies": be, "best_states": bs, "final_states": fs, "best_state": best,
                "best_energy": e, "stats": st.as_dict()}


def _qubo_lib():
    L = lib()
    if not getattr(L, "_qubo_ready", False):
        i64 = ctypes.c_int64
        L.os_qubo_parse.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.os_qubo_parse_device.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.os_qubo_read_binary.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.POINTER(_vp)]
        L.os_qubo_destroy.argtypes = [_vp]
        L.os_qubo_info.argtypes = [_vp, _i32p, _i32p, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.os_qubo_parse_times.argtypes = [_vp, _f32p]
        L.os_qubo_terms.argtypes = [_vp, _i32p, _i32p, _f64p]
        L.os_qubo_flatten.argtypes = [_vp, _f64p]
        L.os_qubo_dense.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.os_qubo_binary_size.argtypes = [_vp]
        L.os_qubo_binary_size.restype = i64
        L.os_qubo_write_binary.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t]
        L._qubo_ready = True
    return L


class DeviceBytes:
    """Bytes copied into GPU memory through the HIP runtime the library itself
    loads (the device of `ctx`), e.g. `.qubo` text for Qubo(ctx, device=...)."""

    def __init__(self, ctx: Context, data):
        lib()
        self._hip = ctypes.CDLL("libamdhip64.so.7")
        data = bytes(data)
        self.size = len(data)
        self.ptr = _vp()
        if self._hip.hipSetDevice(ctypes.c_int(ctx.device)) != 0:
            raise OneSolverError(-2, "hipSetDevice failed")
        if self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.size, 16))) != 0:
            raise OneSolverError(-2, "hipMalloc failed")
        if self.size and self._hip.hipMemcpy(self.ptr, data, ctypes.c_size_t(self.size),
                                             ctypes.c_int(1)) != 0:
            raise OneSolverError(-2, "hipMemcpy failed")
        ctx._adopt(self)

    def close(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def source(self):
        return (self.ptr.value, self.size)


class Qubo:
    """A `.qubo` model parsed on the GPU (os_qubo): the grammar and checks of
    qubo::QUBOModel::load (include/model/qubo.hpp:282-430).  Rejected input
    raises InvalidArgument carrying the reference's exception message."""

    def __init__(self, ctx: Context, text=None, *, binary=None, device=None):
        if sum(x is not None for x in (text, binary, device)) != 1:
            raise ValueError("give one of text, binary, device")
        L = _qubo_lib()
        h = _vp()
        if device is not None:
            ptr, length = device   # text already in GPU memory (16-byte aligned)
            _check(L.os_qubo_parse_device(ctx.handle, ptr, length, ctypes.byref(h)))
        elif text is not None:
            data = text.encode() if isinstance(text, str) else bytes(text)
            _check(L.os_qubo_parse(ctx.handle, data, len(data), ctypes.byref(h)))
        else:
            data = bytes(binary)
            _check(L.os_qubo_read_binary(ctx.handle, data, len(data), ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        ctx._adopt(self)
        n, mx = ctypes.c_int32(), ctypes.c_int32()
        nl, nq = ctypes.c_int64(), ctypes.c_int64()
        _check(L.os_qubo_info(h, ctypes.byref(n), ctypes.byref(mx), ctypes.byref(nl),
                              ctypes.byref(nq)))
        self.n, self.max_nodes = n.value, mx.value
        self.num_linear, self.num_quadratic = nl.value, nq.value

    @classmethod
    def load(cls, ctx: Context, path):
        """Text or binary file, told apart by the binary magic "OSQB"."""
        with open(path, "rb") as f:
            data = f.read()
        if data[:4] == b"OSQB":
            return cls(ctx, binary=data)
        return cls(ctx, data)

    def close(self):
        if self._h:
            _qubo_lib().os_qubo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def parse_times(self):
        """GPU ms of the last parse: line index, line parse, dedupe + compaction."""
        ms = np.zeros(3, np.float32)
        _check(_qubo_lib().os_qubo_parse_times(self._h, _p(ms, _f32p)))
        return ms

    def terms(self):
        """(i, j, v): the distinct terms in order of first appearance, i <= j."""
        T = self.num_linear + self.num_quadratic
        i = np.empty(T, np.int32)
        j = np.empty(T, np.int32)
        v = np.empty(T, np.float64)
        _check(_qubo_lib().os_qubo_terms(self._h, _p(i, _i32p), _p(j, _i32p), _p(v, _f64p)))
        return i, j, v

    def flatten(self):
        """helpers::flatten_qubo: the n x n FP64 symmetric matrix."""
        q = np.empty((self.n, self.n), np.float64)
        _check(_qubo_lib().os_qubo_flatten(self._h, _p(q, _f64p)))
        return q

    def to_binary(self) -> bytes:
        L = _qubo_lib()
        size = L.os_qubo_binary_size(self._h)
        buf = ctypes.create_string_buffer(size)
        _check(L.os_qubo_write_binary(self._h, buf, size))
        return buf.raw

    def dense(self) -> "DenseProblem":
        """A DenseProblem built on the GPU from the terms (os_qubo_dense)."""
        h = _vp()
        _check(_qubo_lib().os_qubo_dense(self._h, ctypes.byref(h)))
        d = DenseProblem.__new__(DenseProblem)
        d.n, d.ctx, d._h = self.n, self.ctx, h
        return d


class CsrParams(ctypes.Structure):
    """os_csr_params of include/onesolver.h."""
    _fields_ = [("num_replicas", ctypes.c_uint32), ("num_sweeps", ctypes.c_int32),
                ("tscale", _vp), ("seed", ctypes.c_uint64), ("instance_base", ctypes.c_uint32),
                ("init_states", _vp), ("use_graph", ctypes.c_int32), ("tune", ctypes.c_int32),
                ("site_order", ctypes.c_int32), ("init_marginals", _vp)]


class CsrPopulationParams(ctypes.Structure):
    """os_csr_population_params of include/onesolver.h."""
    _fields_ = [("num_replicas", ctypes.c_uint32), ("num_steps", ctypes.c_int32),
                ("sweeps_per_step", ctypes.c_int32), ("tscale", _vp), ("seed", ctypes.c_uint64),
                ("instance_base", ctypes.c_uint32), ("site_order", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("tune", ctypes.c_int32)]


class CsrTemperingParams(ctypes.Structure):
    """os_csr_tempering_params of include/onesolver.h."""
    _fields_ = [("num_ladders", ctypes.c_uint32), ("num_temperatures", ctypes.c_uint32),
                ("tscale", _vp), ("num_sweeps", ctypes.c_int32),
                ("exchange_interval", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("instance_base", ctypes.c_uint32), ("site_order", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("tune", ctypes.c_int32)]


def _csr_lib():
    L = lib()
    if not getattr(L, "_csr_ready", False):
        L.os_csr_batch_create.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p,
                                          ctypes.POINTER(_vp)]
        L.os_csr_batch_destroy.argtypes = [_vp]
        L.os_csr_batch_anneal.argtypes = [_vp, ctypes.POINTER(CsrParams), _f64p, _f64p, _u8p,
                                          _u8p, _f64p, _i32p, ctypes.POINTER(SweepStats)]
        L.os_csr_batch_population.argtypes = [_vp, ctypes.POINTER(CsrPopulationParams), _f64p,
                                              _u8p, _u8p, ctypes.POINTER(ctypes.c_uint32),
                                              _f64p, _i32p, ctypes.POINTER(SweepStats)]
        L.os_csr_batch_tempering.argtypes = [_vp, ctypes.POINTER(CsrTemperingParams), _f64p,
                                             _u8p, _u8p, ctypes.POINTER(ctypes.c_uint32),
                                             _f64p, _i32p, ctypes.POINTER(SweepStats)]
        L.os_csr_batch_launch_shape.argtypes = [_vp, _i32p, _i32p, _i32p]
        L.os_csr_batch_colour_shape.argtypes = [_vp, _i32p, _i32p, _i32p]
        L.os_csr_batch_colour_sites.argtypes = [_vp, _i32p]
        L.os_csr_batch_site_order.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, _i32p]
        L._csr_ready = True
    return L


class CsrBatch:
    """A batch of sparse QUBOs resident on the GPU (os_csr_batch).

    ``instances``: sequence of (row_ptr, col, val, diag) with the symmetric
    off-diagonal couplings in CSR (both (j,l) and (l,j)) and the diagonal."""

    def __init__(self, ctx: Context, instances):
        self.ctx = ctx
        self.ns = np.array([len(d) for (_, _, _, d) in instances], dtype=np.int32)
        rp = np.concatenate([np.asarray(r, dtype=np.int32) for (r, _, _, _) in instances])
        cl = np.concatenate([np.asarray(c, dtype=np.int32) for (_, c, _, _) in instances])
        vl = np.concatenate([np.asarray(v, dtype=np.float64) for (_, _, v, _) in instances])
        dg = np.concatenate([np.asarray(d, dtype=np.float64) for (_, _, _, d) in instances])
        rp, cl, vl, dg = (np.ascontiguousarray(a) for a in (rp, cl, vl, dg))
        h = _vp()
        _check(_csr_lib().os_csr_batch_create(ctx.handle, len(self.ns), _p(self.ns, _i32p),
                                              _p(rp, _i32p), _p(cl, _i32p), _p(vl, _f64p),
                                              _p(dg, _f64p), ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    def close(self):
        if self._h:
            _csr_lib().os_csr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def launch_shape(self):
        """(waves per work-group, resident work-groups, 1 if the
        row-prefetching kernel runs else 0); zeros before the first run."""
        w, g, r = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _check(_csr_lib().os_csr_batch_launch_shape(self._h, ctypes.byref(w), ctypes.byref(g),
                                                    ctypes.byref(r)))
        return w.value, g.value, r.value

    def colour_shape(self):
        """(waves per work-group, resident work-groups, sites per step) of the
        colour-class order -- sites per step 0 when the instances differ
        (colour_sites()); zeros before its first run."""
        w, g, s = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _check(_csr_lib().os_csr_batch_colour_shape(self._h, ctypes.byref(w), ctypes.byref(g),
                                                    ctypes.byref(s)))
        return w.value, g.value, s.value

    def colour_sites(self):
        """Sites per wave step of every instance in the colour-class order."""
        out = np.zeros(len(self.ns), dtype=np.int32)
        _check(_csr_lib().os_csr_batch_colour_sites(self._h, _p(out, _i32p)))
        return out

    def site_order(self, instance):
        """(order [n], class starts) of the colour-class order of an instance."""
        n = int(self.ns[instance])
        k = ctypes.c_int32(0)
        _check(_csr_lib().os_csr_batch_site_order(self._h, instance, None, None, ctypes.byref(k)))
        order = np.zeros(n, dtype=np.int32)
        cs = np.zeros(k.value + 1, dtype=np.int32)
        _check(_csr_lib().os_csr_batch_site_order(self._h, instance, _p(order, _i32p),
                                                  _p(cs, _i32p), None))
        return order, cs

    def anneal(self, num_replicas, tscale, seed=1, instance_base=0, init_states=None,
               best_states=True, final_states=False, use_graph=True, tune=True, site_order=0,
               init_marginals=None):
        """Returns per-instance lists of best/final states ([R, n_i] uint8) and
        [I, R] arrays of FP64 best energies and running-best energies.
        site_order 1: sweeps visit the sites class by class (site_order()).
        init_marginals: per-instance [n_i] probabilities, sampled on the GPU."""
        ts = np.ascontiguousarray(tscale, dtype=np.float32)
        I, R = len(self.ns), int(num_replicas)
        ini = None
        if init_states is not None:
            ini = np.ascontiguousarray(np.concatenate(
                [np.asarray(s, dtype=np.uint8).reshape(R, n).ravel()
                 for s, n in zip(init_states, self.ns)]))
        mar = None
        if init_marginals is not None:
            mar = np.ascontiguousarray(np.concatenate(
                [np.asarray(m, dtype=np.float32).reshape(n) for m, n in zip(init_marginals,
                                                                            self.ns)]))
        prm = CsrParams(R, len(ts), ts.ctypes.data, seed, instance_base,
                        None if ini is None else ini.ctypes.data, int(use_graph), int(tune),
                        int(site_order), None if mar is None else mar.ctypes.data)
        tot = int(self.ns.sum()) * R
        be = np.zeros((I, R), dtype=np.float64)
        rb = np.zeros((I, R), dtype=np.float64)
        bs = np.zeros(tot, dtype=np.uint8) if best_states else None
        fs = np.zeros(tot, dtype=np.uint8) if final_states else None
        ib = np.zeros(I, dtype=np.float64)
        ir = np.zeros(I, dtype=np.int32)
        st = SweepStats()
        _check(_csr_lib().os_csr_batch_anneal(self._h, ctypes.byref(prm), _p(be, _f64p),
                                              _p(rb, _f64p), _p(bs, _u8p), _p(fs, _u8p),
                                              _p(ib, _f64p), _p(ir, _i32p), ctypes.byref(st)))

        return {"best_energies": be, "run_best_energies": rb, "best_states": self._split(bs, R),
                "final_states": self._split(fs, R), "inst_best_energy": ib,
                "inst_best_replica": ir, "stats": st.as_dict()}

    def _split(self, flat, R):
        """Per-instance [R, n_i] views of an instance-major state array."""
        if flat is None:
            return None
        out, at = [], 0
        for n in self.ns:
            out.append(flat[at:at + R * n].reshape(R, n))
            at += R * n
        return out

    def tempering(self, ladders, tscale, num_sweeps, exchange_interval, seed=1, instance_base=0,
                  best_states=True, final_states=False, use_graph=True, tune=True,
                  site_order=0):
        """Replica exchange (os_csr_batch_tempering): `ladders` ladders of
        len(tscale) slots per instance; final_slots [I, R]."""
        ts = np.ascontiguousarray(tscale, dtype=np.float32)
        I, K = len(self.ns), len(ts)
        R = int(ladders) * K
        prm = CsrTemperingParams(ladders, K, ts.ctypes.data, num_sweeps, exchange_interval, seed,
                                 instance_base, int(site_order), int(use_graph), int(tune))
        tot = int(self.ns.sum()) * R
        be = np.zeros((I, R), dtype=np.float64)
        bs = np.zeros(tot, dtype=np.uint8) if best_states else None
        fs = np.zeros(tot, dtype=np.uint8) if final_states else None
        sl = np.zeros((I, R), dtype=np.uint32)
        ib = np.zeros(I, dtype=np.float64)
        ir = np.zeros(I, dtype=np.int32)
        st = SweepStats()
        _check(_csr_lib().os_csr_batch_tempering(
            self._h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p), _p(fs, _u8p),
            sl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _p(ib, _f64p), _p(ir, _i32p),
            ctypes.byref(st)))
        return {"best_energies": be, "best_states": self._split(bs, R),
                "final_states": self._split(fs, R), "final_slots": sl, "inst_best_energy": ib,
                "inst_best_replica": ir, "stats": st.as_dict()}

    def population(self, num_replicas, tscale, sweeps_per_step, seed=1, instance_base=0,
                   best_states=True, final_states=False, use_graph=True, tune=True,
                   site_order=0):
        """Population annealing (os_csr_batch_population) through the
        len(tscale) temperature steps; ancestors [I, steps - 1, R]."""
        ts = np.ascontiguousarray(tscale, dtype=np.float32)
        I, R, K = len(self.ns), int(num_replicas), len(ts)
        prm = CsrPopulationParams(R, K, sweeps_per_step, ts.ctypes.data, seed, instance_base,
                                  int(site_order), int(use_graph), int(tune))
        tot = int(self.ns.sum()) * R
        be = np.zeros((I, R), dtype=np.float64)
        bs = np.zeros(tot, dtype=np.uint8) if best_states else None
        fs = np.zeros(tot, dtype=np.uint8) if final_states else None
        anc = np.zeros((I, max(K - 1, 0), R), dtype=np.uint32)
        ib = np.zeros(I, dtype=np.float64)
        ir = np.zeros(I, dtype=np.int32)
        st = SweepStats()
        _check(_csr_lib().os_csr_batch_population(
            self._h, ctypes.byref(prm), _p(be, _f64p), _p(bs, _u8p), _p(fs, _u8p),
            anc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _p(ib, _f64p), _p(ir, _i32p),
            ctypes.byref(st)))
        return {"best_energies": be, "best_states": self._split(bs, R),
                "final_states": self._split(fs, R), "ancestors": anc, "inst_best_energy": ib,
                "inst_best_replica": ir, "stats": st.as_dict()}


def _mpnn_lib():
    L = lib()
    if not getattr(L, "_mpnn_ready", False):
        L.os_mpnn_num_params.restype = ctypes.c_int
        L.os_mpnn_create.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.os_mpnn_destroy.argtypes = [_vp]
        L.os_mpnn_get_params.argtypes = [_vp, _f32p]
        L.os_mpnn_set_params.argtypes = [_vp, _f32p]
        L.os_mpnn_predict.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _f32p]
        L.os_mpnn_predict_csr.argtypes = [_vp, _vp, _vp, _f32p]
        L.os_mpnn_loss_grad.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _u8p,
                                        ctypes.POINTER(ctypes.c_double), _f32p]
        L.os_mpnn_train_step.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _u8p,
                                         ctypes.c_float, ctypes.POINTER(ctypes.c_double)]
        L._mpnn_ready = True
    return L


class WarmStartNet:
    """The warm-start message-passing network on the GPU (os_mpnn)."""

    def __init__(self, ctx: Context, seed: int = 0):
        L = _mpnn_lib()
        self.ctx = ctx
        self.num_params = int(L.os_mpnn_num_params())
        h = _vp()
        _check(L.os_mpnn_create(ctx.handle, seed, ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    def close(self):
        if self._h:
            _mpnn_lib().os_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.num_params, dtype=np.float32)
        _check(_mpnn_lib().os_mpnn_get_params(self._h, _p(out, _f32p)))
        return out

    def set_params(self, flat):
        w = np.ascontiguousarray(flat, dtype=np.float32)
        if w.size != self.num_params:
            raise ValueError("wrong parameter count")
        _check(_mpnn_lib().os_mpnn_set_params(self._h, _p(w, _f32p)))

    def predict(self, data: "DenseBatch", first=0, count=None):
        count = data.I - first if count is None else count
        out = np.zeros((count, data.n), dtype=np.float32)
        _check(_mpnn_lib().os_mpnn_predict(self._h, data._h, first, count, _p(out, _f32p)))
        return out

    def predict_csr(self, batch: "CsrBatch", scale_rows=None):
        """Per-instance marginals [n_i] on a sparse batch (os_mpnn_predict_csr);
        scale_rows: per-instance row count of the scale (the dense layout's
        size), None for n_i."""
        flat = np.zeros(int(batch.ns.sum()), dtype=np.float32)
        rows = None
        if scale_rows is not None:
            rows = np.ascontiguousarray(scale_rows, dtype=np.int32)
            if rows.size != len(batch.ns):
                raise ValueError("scale_rows needs one entry per instance")
        _check(_mpnn_lib().os_mpnn_predict_csr(self._h, batch._h,
                                               None if rows is None else rows.ctypes.data,
                                               _p(flat, _f32p)))
        out, at = [], 0
        for n in batch.ns:
            out.append(flat[at:at + n].copy())
            at += n
        return out

    def loss_grad(self, data: "DenseBatch", labels, first=0):
        y = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1, data.n)
        loss = ctypes.c_double(0.0)
        g = np.zeros(self.num_params, dtype=np.float32)
        _check(_mpnn_lib().os_mpnn_loss_grad(self._h, data._h, first, y.shape[0], _p(y, _u8p),
                                             ctypes.byref(loss), _p(g, _f32p)))
        return loss.value, g

    def train_step(self, data: "DenseBatch", labels, lr, first=0):
        y = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1, data.n)
        loss = ctypes.c_double(0.0)
        _check(_mpnn_lib().os_mpnn_train_step(self._h, data._h, first, y.shape[0], _p(y, _u8p),
                                              lr, ctypes.byref(loss)))
        return loss.value
