"""The allocation-free gradient entry point (nsgpu_mlp_loss_grad_noalloc) against the allocating
one (nsgpu_mlp_loss_grad): same kernels and summation order, so the gradient and the loss must
be bitwise the same floats; two calls in a row, the second with fewer samples (the scratch keeps
the first call's partials), on the null stream and on a stream of the caller; sizes beyond the
scratch are refused before any launch.  Device buffers come from the HIP runtime the library
itself is linked against (loaded by its soname), never from the library.
"""
import ctypes

import numpy as np
import pytest

from conftest import package

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs, on the library's own runtime."""

    def __init__(self):
        package().load_library()  # loads libamdhip64 as a dependency
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n):
        out = np.zeros(n, dtype=np.float32)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 8)).astype(np.float32)
    Q = rng.standard_normal((n, 3)).astype(np.float32)
    T = (Q.mean(axis=1) + 0.1 * rng.standard_normal(n)).astype(np.float32)
    return X, Q, T


@pytest.fixture(scope="module")
def params():
    import importlib
    return importlib.import_module("3dnssolver_amd.learned").load()


def test_noalloc_gradient_equals_the_allocating_one(params):
    ns = package()
    L = ns.load_library()
    hip = Hip()
    stream = ctypes.c_void_p()
    assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
    P = ns.MLP_NPARAM
    try:
        dparams = hip.put(params)
        for n, seed, st in ((200_000, 1, None), (5_000, 2, stream), (ns_max(), 3, stream)):
            print(f"[gpu] noalloc gradient n={n}", flush=True)
            X, Q, T = _samples(n, seed)
            loss, grad = ns.mlp_loss_grad(params, X, Q, T)
            dX, dQ, dT = hip.put(X), hip.put(Q), hip.put(T)
            dout = hip.put(np.full(P + 1, np.nan, dtype=np.float32))
            rc = L.nsgpu_mlp_loss_grad_noalloc(0, dparams, n, dX, dQ, dT, dout, st)
            assert rc == 0
            assert hip.L.hipStreamSynchronize(st) == 0
            out = hip.get(dout, P + 1)
            np.testing.assert_array_equal(out[:P], grad)
            assert np.float32(out[P]) == np.float32(loss)
        # beyond the scratch: refused before any launch
        rc = L.nsgpu_mlp_loss_grad_noalloc(0, dparams, ns_max() + 1, dX, dQ, dT, dout, None)
        assert rc == -1
        rc = L.nsgpu_mlp_loss_grad_noalloc(0, dparams, 0, dX, dQ, dT, dout, None)
        assert rc == -1
    finally:
        hip.free()
        hip.L.hipStreamDestroy(stream)


def ns_max():
    return 262144  # NSGPU_MLP_NOALLOC_MAX_N
