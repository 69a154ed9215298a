"""dcg_mlp_eval_device, the entry point that allocates nothing, against
dcg_mlp_eval (host buffers, allocating): two calls in a row, the second with fewer
rows, outputs and gradients equal bit for bit (the same passes, the same grid and
the same reduction order).  The test's own device buffers come from the HIP
runtime the library links (hipMalloc through ctypes), so one runtime serves the
process; the calls go through the C ABI."""
import ctypes as C
import os

import numpy as np
import pytest

from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libdcg.so links (same soname, so the same loaded copy)."""
    import dycgraph_amd as d
    d.lib()
    for name in ("libamdhip64.so.7", os.path.join("/opt/rocm/lib", "libamdhip64.so.7"), "libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipSetDevice.argtypes = [C.c_int]
    return h


class DeviceArray:
    H2D, D2H = 1, 2

    def __init__(self, hip, host):
        self.hip, self.shape, self.nbytes = hip, host.shape, host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(host.nbytes, 4)) == 0
        self.put(host)

    def put(self, host):
        a = np.ascontiguousarray(host, np.float32)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, self.H2D) == 0

    def get(self):
        a = np.empty(self.shape, np.float32)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, self.D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def _case(rng, rows):
    shapes = [(8, 64), (64,), (64, 64), (64,), (64, 2), (2,)]
    params = np.concatenate([rng.normal(0, 0.3, s).astype(np.float32).reshape(-1) for s in shapes])
    x = np.abs(rng.normal(0, 2, (rows, 8))).astype(np.float32)
    y = np.stack([rng.integers(0, 2, rows), np.abs(rng.normal(2, 1, rows))], 1).astype(np.float32)
    return x, y, params


def test_no_alloc_mlp_matches_allocating_entry_point():
    import dycgraph_amd as d
    L = d.lib()
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    rng = np.random.default_rng(77)
    for rows in (1000, 100):
        announce(f"no-alloc mlp rows={rows}")
        x, y, params = _case(rng, rows)
        out_ref = np.empty((rows, 2), np.float32)
        grad_ref = np.empty(params.size, np.float32)
        assert L.dcg_mlp_eval(0, rows, x.ctypes.data, y.ctypes.data, params.ctypes.data, out_ref.ctypes.data,
                              grad_ref.ctypes.data) == 0
        nan_out = np.full((rows, 2), np.nan, np.float32)
        dx, dy, dp = DeviceArray(hip, x), DeviceArray(hip, y), DeviceArray(hip, params)
        dout = DeviceArray(hip, nan_out)
        dgrad = DeviceArray(hip, np.full(params.size, np.nan, np.float32))
        assert L.dcg_mlp_eval_device(0, rows, dx.ptr, dy.ptr, dp.ptr, dout.ptr, dgrad.ptr) == 0
        out, grad = dout.get(), dgrad.get()
        assert np.array_equal(out, out_ref), float(np.nanmax(np.abs(out - out_ref)))
        assert np.array_equal(grad, grad_ref), float(np.nanmax(np.abs(grad - grad_ref)))
        # forward only (no targets, no gradient)
        dout.put(nan_out)
        assert L.dcg_mlp_eval_device(0, rows, dx.ptr, None, dp.ptr, dout.ptr, None) == 0
        assert np.array_equal(dout.get(), out_ref)
        for a in (dx, dy, dp, dout, dgrad):
            a.free()


def test_no_alloc_rejects_oversized_and_bad_arguments():
    import dycgraph_amd as d
    L = d.lib()
    hip = _hip()
    announce("no-alloc mlp argument checks")
    buf = DeviceArray(hip, np.zeros(16, np.float32))
    p = buf.ptr
    assert L.dcg_mlp_eval_device(0, (1 << 24) + 1, p, None, p, p, None) == -4   # DCG_ERR_CAPACITY
    assert L.dcg_mlp_eval_device(0, 0, p, None, p, p, None) == -1               # DCG_ERR_ARG
    assert L.dcg_mlp_eval_device(0, 4, p, None, p, p, p) == -1                  # grad without targets
    buf.free()
