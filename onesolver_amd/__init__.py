"""Python binding of libonesolver.so, the MI355X simulated-annealing QUBO library.

Thin ctypes layer over the C ABI declared in ``include/onesolver.h``; used by the
tests and the benchmark.  It never falls back to a CPU implementation: if the
library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libonesolver.so")
CLI_PATH = os.path.join(HERE, "bin", "one-solver-anneal")
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
        _lib = L
    return _lib


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
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
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
