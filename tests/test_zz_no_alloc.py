"""pib_tridiag_batched_noalloc: the cyclic tridiagonal drop-in on device
memory with its workspace in the library's static device array.  Checked
against pib_tridiag_batched (same kernel, host buffers, allocating) bit for
bit, on two calls in a row, the second smaller, on a stream of the caller's.
Device memory and the stream come from the HIP runtime the library links
(libamdhip64 through ctypes), so no second runtime is loaded."""
import ctypes

import numpy as np
import pytest

from pureibm_ps2d_fd1d_amd import _abi

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs, from the runtime the HIP
    library itself links."""

    def __init__(self):
        _abi.load_gpu()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        self.h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        self.h.hipFree.argtypes = [vp]
        self.h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        self.h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.h.hipStreamSynchronize.argtypes = [vp]
        self.h.hipStreamDestroy.argtypes = [vp]
        self.bufs = []

    def ok(self, e):
        assert e == 0, f"HIP error {e}"

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(d), a.nbytes))
        self.bufs.append(d)
        self.ok(self.h.hipMemcpy(d, a.ctypes.data, a.nbytes, 1))  # host to device
        return d

    def get(self, d, shape):
        a = np.zeros(shape)
        self.ok(self.h.hipMemcpy(a.ctypes.data, d, a.nbytes, 2))  # device to host
        return a

    def free(self):
        for d in self.bufs:
            self.h.hipFree(d)
        self.bufs = []


def _system(rng, n, ns):
    a = rng.uniform(0.2, 1.0, (n, ns))
    c = rng.uniform(0.2, 1.0, (n, ns))
    b = -2.5 - rng.uniform(0.0, 1.0, (n, ns))
    r = rng.standard_normal((n, ns))
    return a, b, c, r


def _ctx():
    c = _abi.PibConfig()
    c.nx, c.my, c.mz, c.precision = 8, 8, 8, 0
    c.dx, c.vis, c.cfl_max = 0.1, 0.01, 0.5
    c.ucharmod, c.char_length, c.dia, c.dr, c.voldom = 1.0, 1.0, 1.0, 1.0, 1.0
    lib = _abi.load_gpu()
    ctx = ctypes.c_void_p()
    assert lib.pib_create(ctypes.byref(c), ctypes.byref(ctx)) == 0
    return lib, ctx


@pytest.mark.parametrize("cyclic", [1, 0])
def test_noalloc_tridiag_matches_the_allocating_entry_point(cyclic):
    lib, ctx = _ctx()
    hip = Hip()
    stream = ctypes.c_void_p()
    hip.ok(hip.h.hipStreamCreate(ctypes.byref(stream)))
    rng = np.random.default_rng(40 + cyclic)
    d = _abi.dptr
    # the 8 320 modes of a 128^3 grid over 126 rows, then a smaller call
    for n, ns in ((126, 8320), (37, 501)):
        print(f"noalloc tridiag n={n} nsys={ns} cyclic={cyclic}", flush=True)
        a, b, c, r = _system(rng, n, ns)
        ref = np.zeros((n, ns))
        assert lib.pib_tridiag_batched(ctx, n, ns, d(a), d(b), d(c), d(r), d(ref), cyclic) == 0
        dv = [hip.put(v) for v in (a, b, c, r)]
        dx = hip.put(np.full((n, ns), np.nan))
        assert lib.pib_tridiag_batched_noalloc(ctx, n, ns, *dv, dx, cyclic, stream) == 0
        hip.ok(hip.h.hipStreamSynchronize(stream))
        x = hip.get(dx, (n, ns))
        assert np.array_equal(x, ref)
        # the solution solves the system (the comparison is not vacuous)
        res = b * x + a * np.roll(x, 1, axis=0) + c * np.roll(x, -1, axis=0) - r
        if not cyclic:
            res[0] -= a[0] * x[-1]
            res[-1] -= c[-1] * x[0]
        assert np.max(np.abs(res)) < 1e-12 * np.max(np.abs(r))
        hip.free()
    hip.ok(hip.h.hipStreamDestroy(stream))
    lib.pib_destroy(ctx)


def test_noalloc_tridiag_rejects_sizes_beyond_the_workspace():
    lib, ctx = _ctx()
    n = 3
    ns = _abi.PIB_TRIDIAG_NOALLOC_MAX // n + 1
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    assert lib.pib_tridiag_batched_noalloc(ctx, n, ns, fake, fake, fake, fake, fake, 1, None) == _abi.PIB_ERR_ARG
    assert b"PIB_TRIDIAG_NOALLOC_MAX" in lib.pib_last_error(ctx)
    assert lib.pib_tridiag_batched_noalloc(None, 2, 4, fake, fake, fake, fake, fake, 1, None) == _abi.PIB_ERR_ARG
    lib.pib_destroy(ctx)
