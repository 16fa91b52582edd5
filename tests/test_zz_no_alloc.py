"""pdhg_dev_evaluate_noalloc (include/pdhg_hip.h): the evaluate() of pdhg_dev_evaluate with
device pointers and the library's per-GPU scratch array instead of the handle's buffers. Its
output must equal pdhg_dev_evaluate's bit for bit (the same kernels walk the same work items
and add the same partial sums in the same order), on two calls in a row, the second on a
smaller LP, so that the second call runs over scratch the first one left behind."""
import ctypes as C

import numpy as np
import pytest

import common as cm
import sankhya_hip as sh


def _hip():
    """HIP's runtime as the engine library bound it. The engine library is loaded first; then
    the runtime is opened by its soname, which the dynamic loader resolves to the copy already
    loaded under that name, the same rule that bound the engine library's own dependency: ROCm's
    runtime when the engine came first, torch's bundled copy when torch was imported earlier in
    the process (opening ROCm's by its path then would load a second runtime against torch's
    older HSA library, and fail)."""
    sh.load_library()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _noalloc(dev, x, y):
    lib = sh.load_library()
    lib.pdhg_dev_evaluate_noalloc.argtypes = [C.c_void_p] * 6
    lib.pdhg_dev_evaluate_noalloc.restype = C.c_int
    hip = _hip()
    m, n = dev.m, dev.n
    host = {"x": np.ascontiguousarray(x, np.float64), "y": np.ascontiguousarray(y, np.float64),
            "act": np.full(max(m, 1), np.nan), "red": np.full(max(n, 1), np.nan),
            "res": np.full(9, np.nan)}
    dptr = {}
    try:
        for k, a in host.items():
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
            dptr[k] = p
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        rc = lib.pdhg_dev_evaluate_noalloc(dev._h, dptr["x"], dptr["y"], dptr["act"], dptr["red"],
                                           dptr["res"])
        assert rc == 0, lib.pdhg_dev_last_error(dev._h)
        for k in ("act", "red", "res"):
            assert hip.hipMemcpy(host[k].ctypes.data, dptr[k], host[k].nbytes, D2H) == 0
    finally:
        for p in dptr.values():
            hip.hipFree(p)
    return host["act"][:m], host["red"][:n], host["res"]


@pytest.mark.gpu
def test_noalloc_evaluate_equals_evaluate_on_two_calls():
    cm.gpu_or_skip()
    cases = [("zipf-20000x50000 (long rows)", cm.generated("zipf", 20000, 50000, 10, 3)[0]),
             ("c0 (smaller)", cm.c0(2)[0])]
    rng = np.random.default_rng(11)
    for name, mdl in cases:
        m, n, nnz = mdl.dims
        dev = sh.Device(mdl)
        x = rng.uniform(-1.0, 11.0, n)
        y = rng.standard_normal(m)
        a0, r0, s0 = dev.evaluate(x, y)
        a1, r1, s1 = _noalloc(dev, x, y)
        dev.close()
        print(f"{name}: m={m} n={nnz and n} residuals {s1[:3]}", flush=True)
        assert np.array_equal(a0, a1) and np.array_equal(r0, r1) and np.array_equal(s0, s1), name
