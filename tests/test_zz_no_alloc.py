"""os_anneal_reference_noalloc through the C ABI against os_anneal_reference.

The no-allocation entry point runs the same chains as os_anneal_reference with
its working arrays in a fixed per-GPU scratch array; inputs and outputs are
device arrays, which this test allocates with the HIP runtime the library
loads.  Two calls in a row, the second on a smaller instance with fewer
tries, must each equal os_anneal_reference's states, energies and best index
exactly, and a call must not lower the device's free memory.
"""
import ctypes
import os

import numpy as np
import pytest

import onesolver_amd
import oracle
from conftest import GOLDEN
from qubo_io import load_qubo

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


class _Hip:
    def __init__(self, device):
        onesolver_amd.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        assert self.h.hipSetDevice(ctypes.c_int(device)) == 0
        self.bufs = []

    def alloc(self, nbytes):
        p = _vp()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 16))) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(_vp), ctypes.c_size_t(a.nbytes),
                                ctypes.c_int(1)) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(_vp), p, ctypes.c_size_t(out.nbytes),
                                ctypes.c_int(2)) == 0
        return out

    def free_bytes(self):
        free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return free.value

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(ctx, hip, q, beta, num_tries, seed):
    L = onesolver_amd.lib()
    L.os_anneal_reference_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                              _vp, _vp, _vp]
    n = q.shape[0]
    dq = hip.upload(np.ascontiguousarray(q, dtype=np.float64))
    db = hip.upload(np.ascontiguousarray(beta, dtype=np.float64))
    ds = hip.alloc(num_tries * n)
    de = hip.alloc(8 * num_tries)
    di = hip.alloc(4)
    free0 = hip.free_bytes()
    rc = L.os_anneal_reference_noalloc(ctx.handle, n, dq, db, len(beta), 1, num_tries, seed,
                                       ds, de, di)
    free1 = hip.free_bytes()
    assert rc == 0, L.os_last_error()
    assert free1 >= free0, (free0, free1)
    return (hip.download(ds, (num_tries, n), np.uint8), hip.download(de, num_tries, np.float64),
            int(hip.download(di, 1, np.int32)[0]))


def test_noalloc_equals_the_allocating_entry_point_on_two_calls(ctx):
    L = onesolver_amd.lib()
    L.os_noalloc_scratch_bytes.restype = ctypes.c_size_t
    assert L.os_noalloc_scratch_bytes() == 16 << 20
    hip = _Hip(ctx.device)
    try:
        cases = [
            (load_qubo(os.path.join(GOLDEN, "chimera", "128_001.qubo")),
             oracle.geometric_schedule(0.1, 10.0, 1000), 2048, 77),
            (load_qubo(os.path.join(GOLDEN, "c1_dense28.qubo")),
             oracle.geometric_schedule(0.1, 1.0, 300), 700, 1234),
        ]
        for q, beta, tries, seed in cases:
            n = q.shape[0]
            print(f"[gpu] noalloc n={n} tries={tries}", flush=True)
            best, be, states, energies = ctx.anneal_reference(q, beta, tries, seed=seed)
            s2, e2, i2 = _noalloc(ctx, hip, q, beta, tries, seed)
            assert (s2 == states).all()
            assert np.array_equal(e2, energies)
            assert i2 == int(np.argmin(energies)) and e2[i2] == be and (s2[i2] == best).all()
    finally:
        hip.close()


def test_noalloc_rejects_runs_beyond_its_scratch(ctx):
    L = onesolver_amd.lib()
    L.os_anneal_reference_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                              _vp, _vp, _vp]
    hip = _Hip(ctx.device)
    try:
        dq = hip.alloc(8)
        # 2048 spins x 65536 tries: 64 * 256 + 16 + 8 * 32 bytes per try > 16 MiB
        rc = L.os_anneal_reference_noalloc(ctx.handle, 2048, dq, None, 0, 1, 65536, 1, None,
                                           None, None)
        assert rc == onesolver_amd.OS_ERR_UNSUPPORTED
        assert b"scratch" in L.os_last_error()
    finally:
        hip.close()
