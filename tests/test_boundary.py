"""The C-ABI library: it loads without a GPU, exports every entry point that
include/hydra_hip.h declares, and rejects bad input the way the reference
does (an invalid flux choice stops the run, mod_Fluxmodule.F90:14-16)
before it touches the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from hydraphase import _lib
from hydraphase import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hydra_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void|size_t)\s*\*?\s*(hp_\w+)\s*\(", txt, re.M)))


def test_header_declares_all_bound_symbols():
    names = declared()
    assert "hp_create" in names and "hp_step" in names
    assert sorted(_lib.EXPORTS) == names


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    for n in declared():
        assert hasattr(L, n), n


def _create(mesh, params):
    L = _lib.lib()
    ne, ng = mesh.neles, mesh.nghosts
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    nod, nc = i32(mesh.nod + 1), i32(mesh.nc + 1)
    ib, gpar, gid = i32(mesh.iblank), i32(mesh.gparent + 1), i32(np.arange(ng) + ne + 1)
    gt, gs = i32(mesh.gtype), i32(mesh.gside)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    p = _lib.make_params(params)
    h = ctypes.c_void_p()
    rc = L.hp_create(ctypes.byref(h), mesh.nodes, ne, ng, dp(mesh.x), dp(mesh.y), dp(mesh.z), ip(nod), ne,
                     ip(nc[0]), ip(nc[1]), ip(nc[2]), ip(nc[3]), ip(ib), ip(gpar), ip(gid), ip(gt), ip(gs),
                     ctypes.byref(p))
    return rc, L.hp_last_error(None).decode(), h


def test_invalid_flux_method_rejected():
    m = M.box_mesh(2, 2, 2)
    p = M.water_air_params(myflux=3)
    rc, msg, h = _create(m, p)
    assert rc == 4 and "Invalid flux method" in msg and not h.value


def test_inconsistent_mesh_rejected():
    m = M.box_mesh(2, 2, 2)
    nc = m.nc.copy()
    nc[0, 0] = m.neles + m.nghosts + 5  # neighbour id out of range
    bad = M.Mesh(m.x, m.y, m.z, m.nod, nc, m.iblank, m.gparent, m.gtype, m.gside, m.shape)
    rc, msg, h = _create(bad, M.water_air_params())
    assert rc == 3 and "neighbour" in msg
    gp = m.gparent.copy()
    gp[0] = (gp[0] + 1) % m.neles  # ghost no longer on its parent's face
    bad = M.Mesh(m.x, m.y, m.z, m.nod, m.nc, m.iblank, gp, m.gtype, m.gside, m.shape)
    rc, msg, h = _create(bad, M.water_air_params())
    assert rc == 3


def test_null_arguments_rejected():
    L = _lib.lib()
    assert L.hp_create(None, 1, 1, 0, None, None, None, None, 1, None, None, None, None, None, None, None,
                       None, None, None) == 1
    assert L.hp_step(None, 1) == 1
    assert L.hp_upload_state(None, None, 0, None, 0.0, 0) == 1


def test_bad_order_and_precision_rejected():
    m = M.box_mesh(2, 2, 2)
    p = M.water_air_params()
    p["myorder"] = 3
    rc, msg, h = _create(m, p)
    assert rc == 1 and "myorder" in msg and not h.value
    p = M.water_air_params()
    p["precision"] = 2
    rc, msg, h = _create(m, p)
    assert rc == 1 and "precision" in msg and not h.value


def test_health_and_division_check_arguments():
    L = _lib.lib()
    st = ctypes.c_int(7)
    assert L.hp_health(None, ctypes.byref(st)) == 1
    assert L.hp_batch_health(None, ctypes.byref(st), 1) == 1
    n = ctypes.c_size_t(3)
    # no pairs: nothing to check, no device work
    assert L.hp_check_division(None, None, 0, ctypes.byref(n)) == 0 and n.value == 0
    assert L.hp_check_division(None, None, 4, ctypes.byref(n)) == 1
    assert "hp_check_division" in L.hp_last_error(None).decode()


def test_dt_min_noalloc_rejects_bad_arguments_before_any_device_call():
    """hp_dt_min_noalloc checks its arguments on the host first: null
    arrays, ldq < neles, neles <= 0 and neles above HP_NOALLOC_MAX_CELLS are
    HP_ERR_ARG with a message (no device is touched, so this runs without a
    GPU)."""
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: rejected before use
    mx = _lib.NOALLOC_MAX_CELLS
    cases = [(None, 10, fake, 10, fake), (fake, 10, None, 10, fake), (fake, 10, fake, 10, None),
             (fake, 5, fake, 10, fake), (fake, 10, fake, 0, fake), (fake, mx + 1, fake, mx + 1, fake)]
    for q, ldq, dl, n, out in cases:
        rc = L.hp_dt_min_noalloc(q, ldq, dl, n, 1.4, 4.4, 6e8, out, None)
        assert rc == 1, (ldq, n, rc)
        assert L.hp_last_error(None)
    assert b"HP_NOALLOC_MAX_CELLS" in L.hp_last_error(None)
