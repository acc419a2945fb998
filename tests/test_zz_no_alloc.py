"""hp_dt_min_noalloc, the CFL min-reduction that allocates nothing, against
hp_dt_min on the same primitives, bit for bit, on two calls in a row (the
second on a smaller case, so the shared scratch array holds the first
call's larger set of partials when the second runs).

The test's own device buffers come from the HIP runtime the library itself
loaded (libamdhip64.so.7, through ctypes), not from torch, which carries a
HIP runtime of its own."""
import ctypes

import numpy as np
import pytest

from hydraphase import mesh as M

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    from hydraphase import _lib
    _lib.lib()  # the library's own libamdhip64.so.7 is now loaded
    h = ctypes.CDLL("libamdhip64.so.7")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
        getattr(h, f).restype = ctypes.c_int
    return h


def to_device(h, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def case_prims(n, seed, steps):
    from hydraphase import Solver
    m = M.box_mesh(n, n, n, jitter=0.1, seed=seed)
    p = M.water_air_params(cfl=0.5)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.05, 0.0, 0.2)])
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.step(steps)
    s.primitive()
    ref = s.dt_min()
    q, dl = s.get("q"), s.get("dl")
    s.close()
    return p, q, dl, ref


def test_dt_min_noalloc_equals_dt_min():
    from hydraphase import _lib
    L = _lib.lib()
    h = hip_runtime()
    for n, seed, steps in ((24, 1, 3), (7, 4, 2)):  # 82,944 cells, then 2,058
        p, q, dl, ref = case_prims(n, seed, steps)
        ne = dl.size
        print(f"noalloc dt: {ne} cells", flush=True)
        dq, ddl = to_device(h, q), to_device(h, dl)
        dout = to_device(h, np.array([-1.0]))
        try:
            rc = L.hp_dt_min_noalloc(dq, q.shape[1], ddl, ne, p["gamag"], p["gamap"], p["pinf"], dout, None)
            assert rc == 0, L.hp_last_error(None)
            assert h.hipDeviceSynchronize() == 0
            got = np.zeros(1)
            assert h.hipMemcpy(got.ctypes.data, dout, 8, D2H) == 0
        finally:
            for ptr in (dq, ddl, dout):
                h.hipFree(ptr)
        assert np.isfinite(ref) and ref > 0
        assert got.view(np.uint64)[0] == np.float64(ref).view(np.uint64), (ne, got[0], ref)
