"""wf_mlp_grad_noalloc (the surrogate's loss gradient on caller-provided
device arrays, partial sums in the library's fixed device array) against
wf_mlp_grad (device memory of a domain handle) on the same inputs.

Both run the same kernels with the same block split, so the gradient and
the loss must be bit-identical.  Two calls in a row, the second with fewer
samples: the fixed array then still holds the first call's partials past
the second call's blocks, which must not leak into its result.  The
caller's device arrays come from the HIP runtime the library itself links
(hipMalloc through ctypes), as a caller's own would."""
import ctypes as C

import numpy as np
import pytest

from weldformgpu_amd import _abi

from harness import small_c1, to_gpu

pytestmark = pytest.mark.gpu

P = 1249
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.restype = C.c_int
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.restype = C.c_int
    return h


class DevBuf:
    """A device array of float32 owned by the test (the caller)."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 4 * max(n, 1)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32).ravel()
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 4 * self.n, H2D) == 0

    def get(self):
        a = np.empty(self.n, np.float32)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, 4 * self.n, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(lib):
    f = lib.wf_mlp_grad_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3
    return f


def _data(seed, n):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.3, P).astype(np.float32)
    X = rng.uniform(0, 1.5, (n, 4)).astype(np.float32)
    Y = rng.uniform(0, 1, n).astype(np.float32)
    return w, X, Y


def test_noalloc_gradient_matches_domain_entry_point(wflib):
    fn = _noalloc(wflib)
    hip = _hip()
    with to_gpu(small_c1(0.3), wflib) as d:
        for seed, n in ((5, 12000), (6, 700)):
            w, X, Y = _data(seed, n)
            print(f"[case] noalloc grad n={n}", flush=True)
            d.set_mlp_params(w)
            g_ref, loss_ref = d.mlp_grad(X, Y)
            bufs = [DevBuf(hip, P), DevBuf(hip, 4 * n), DevBuf(hip, n), DevBuf(hip, P),
                    DevBuf(hip, 1)]
            bw, bX, bY, bg, bl = bufs
            bw.put(w)
            bX.put(X)
            bY.put(Y)
            bg.put(np.full(P, np.nan, np.float32))
            bl.put(np.full(1, np.nan, np.float32))
            rc = fn(bw.p, bX.p, bY.p, n, bg.p, bl.p, None)
            assert rc == 0, wflib.wf_last_error()
            assert hip.hipDeviceSynchronize() == 0
            g, loss = bg.get(), float(bl.get()[0])
            for b in bufs:
                b.free()
            assert np.array_equal(g, g_ref), float(np.abs(g - g_ref).max())
            assert loss == np.float32(loss_ref)


def test_noalloc_rejects_bad_arguments(wflib):
    fn = _noalloc(wflib)
    hip = _hip()
    wflib.wf_last_error.restype = C.c_char_p
    b = DevBuf(hip, 4 * 64 + P + 8)
    p = b.p
    for n in (0, -1, (1 << 24) + 1):
        assert fn(p, p, p, n, p, p, None) == _abi.WF_E_ARG
        assert b"sample count" in wflib.wf_last_error()
    assert fn(None, p, p, 10, p, p, None) == _abi.WF_E_ARG
    assert b"null" in wflib.wf_last_error()
    b.free()
