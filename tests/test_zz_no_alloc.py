"""nphys_mlp_gradients_noalloc (the entry point that allocates nothing: its batch buffers live in a
fixed device array of the library) against nphys_mlp_gradients, through the C ABI: two calls in a
row, the second with fewer rows, must give the same gradients and loss bit for bit (same kernels,
same inputs, same reduction order); more rows than the stated maximum are rejected.

Device buffers for the test come from the HIP runtime the library itself links (found in this
process's mappings), so both see one runtime."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 4096
H2D, D2H = 1, 2


def _hip_runtime():
    for line in open("/proc/self/maps"):
        path = line.split()[-1] if len(line.split()) >= 6 else ""
        if "libamdhip64.so" in path:
            hip = C.CDLL(path)
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipFree.argtypes = [C.c_void_p]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipDeviceSynchronize.argtypes = []
            return hip
    raise RuntimeError("the library's HIP runtime is not mapped")


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 4)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _rows(rng, n):
    X = rng.normal(0.0, 1.0, (n, 24)).astype(np.float32)
    T = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
    return X, T


def test_noalloc_gradients_match_the_allocating_entry_point(nphys):
    from nphysics_amd import learned

    L = nphys.load_library()
    learned._declare(L)
    vp = C.c_void_p
    L.nphys_mlp_gradients_noalloc.argtypes = [vp, vp, vp, C.c_int64, vp, vp, vp]
    net = learned.WarmStartNet(seed=7)  # initialises the device through the library
    hip = _hip_runtime()
    rng = np.random.default_rng(11)
    # normalisation from data, so the scales are not the identity
    Xd, Td = _rows(rng, 2048)
    net.set_data(Xd * 1.7 + 0.3, Td * 0.8)
    xd = DevBuf(hip, (MAX_ROWS + 1) * 24 * 4)
    td = DevBuf(hip, (MAX_ROWS + 1) * 3 * 4)
    gd = DevBuf(hip, learned.NUM_TRAINED * 4)
    ld = DevBuf(hip, 8)
    try:
        for n in (MAX_ROWS, 1000):
            X, T = _rows(rng, n)
            g_ref, loss_ref = net.gradients(X, T)
            xd.put(X)
            td.put(T)
            gd.put(np.full(learned.NUM_TRAINED, np.nan, np.float32))
            ld.put(np.full(1, np.nan))
            print(f"no-alloc gradients, rows {n}", flush=True)
            rc = L.nphys_mlp_gradients_noalloc(net._h, xd.p, td.p, n, gd.p, ld.p, None)
            assert rc == 0, L.nphys_last_error().decode()
            g = gd.get(learned.NUM_TRAINED, np.float32)
            loss = float(ld.get(1, np.float64)[0])
            print(f"rows {n}: loss {loss:.6g} (allocating entry point {loss_ref:.6g}), |g| {np.linalg.norm(g):.6g}")
            assert np.array_equal(g, g_ref), f"rows {n}: max diff {np.max(np.abs(g - g_ref))}"
            assert loss == loss_ref
        rc = L.nphys_mlp_gradients_noalloc(net._h, xd.p, td.p, MAX_ROWS + 1, gd.p, ld.p, None)
        assert rc != 0 and "rows must be" in L.nphys_last_error().decode()
    finally:
        for b in (xd, td, gd, ld):
            b.free()
        net.close()
