"""ctypes binding of the CPU oracle (libhydra_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhydra_oracle.so")

PARAM_D = ["pmax", "pmin", "tgmax", "tgmin", "ugmax", "ugmin", "vgmax", "vgmin", "wgmax", "wgmin",
           "upmax", "upmin", "vpmax", "vpmin", "wpmax", "wpmin", "tpmax", "tpmin",
           "pinf", "gamap", "gamag", "cpp", "cpg", "amolg", "prg",
           "epslnmax", "epslnmin", "epslninf", "phipzero",
           "cfl", "diap", "surfacetension", "as_", "ts"]
PARAM_I = ["niter", "itersub", "istart", "myflux", "myorder", "pad_"]
PARAM_X = ["co2", "co4", "c23", "cmug", "rrg", "gravity", "prp", "foursigmabyd"]


class Params(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in PARAM_D] + [(n, ctypes.c_int) for n in PARAM_I]
                + [(n, ctypes.c_double) for n in PARAM_X])


def make_params(d: dict) -> Params:
    p = Params()
    for n in PARAM_D + PARAM_I:
        setattr(p, n, d.get(n, 0))
    for n in PARAM_X:
        setattr(p, n, d.get(n, 0.0))
    return p


# prim / gradient / array slots (hydra_oracle.h)
Q_NAMES = ["ug", "vg", "wg", "pg", "pp", "tg", "phig", "up", "vp", "wp", "phip", "tp", "rop", "rog",
           "egt", "ept", "ag", "ap"]
G_NAMES = ["ug", "vg", "wg", "tg", "up", "vp", "wp", "tp", "pg", "pp", "rog", "rop", "phig"]
# face_state fields (hyo_face_states)
F_NAMES = ["rog", "ug", "vg", "wg", "tg", "phig", "rop", "up", "vp", "wp", "tp", "pg", "pp", "phip"]
HA = {n: i for i, n in enumerate(["cn", "cu", "cn1", "rhs", "d24", "d2q", "anu", "q", "grad", "phi",
                                   "vis", "sc", "xc", "yc", "zc", "vol", "dl", "dd", "epssq"])}


def build():
    subprocess.run(["make", "-s", "-C", HERE, "libhydra_oracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.hyo_create.restype = vp
        L.hyo_create.argtypes = [ctypes.c_int] * 3 + [dp] * 3 + [ip] * 6 + [ctypes.POINTER(Params)]
        L.hyo_destroy.argtypes = [vp]
        L.hyo_set_state.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.hyo_get_state.argtypes = [vp, dp, dp, dp, ip]
        for n in ["hyo_primitive", "hyo_primitiveplusbc", "hyo_smooth", "hyo_viscousgp",
                  "hyo_viscous_bc", "hyo_bc", "hyo_derivative", "hyo_venkat", "hyo_face_flux",
                  "hyo_source", "hyo_flux_solver", "hyo_rk"]:
            getattr(L, n).argtypes = [vp]
        L.hyo_rk_stage.argtypes = [vp, ctypes.c_int]
        L.hyo_dtmin.argtypes = [vp]
        L.hyo_dtmin.restype = ctypes.c_double
        L.hyo_set_dt.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.hyo_step.argtypes = [vp, dp]
        L.hyo_run.argtypes = [vp, ctypes.c_int]
        L.hyo_array.argtypes = [vp, ctypes.c_int]
        L.hyo_array.restype = ctypes.c_void_p
        L.hyo_get_dt.argtypes = [vp]
        L.hyo_set_donors.argtypes = [vp, ctypes.c_int, ip, ip, dp]
        L.hyo_donor_gather.argtypes = [vp]
        L.hyo_get_dt.restype = ctypes.c_double
        L.hyo_face_states.argtypes = [vp, dp]
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One case on the CPU oracle."""

    def __init__(self, mesh, params: dict):
        L = lib()
        self.mesh = mesh
        self.params = dict(params)
        self._keep = [np.ascontiguousarray(a) for a in
                      (mesh.x, mesh.y, mesh.z)] + [np.ascontiguousarray(a, dtype=np.int32) for a in
                                                   (mesh.nod, mesh.nc, mesh.iblank, mesh.gparent,
                                                    mesh.gtype, mesh.gside)]
        x, y, z, nod, nc, ib, gpar, gt, gs = self._keep
        self._p = make_params(params)
        self.h = L.hyo_create(mesh.nodes, mesh.neles, mesh.nghosts, _dp(x), _dp(y), _dp(z),
                              _ip(nod), _ip(nc), _ip(ib), _ip(gpar), _ip(gt), _ip(gs),
                              ctypes.byref(self._p))
        self.ne, self.ng, self.nt = mesh.neles, mesh.nghosts, mesh.ntot

    def __del__(self):
        if getattr(self, "h", None):
            lib().hyo_destroy(self.h)
            self.h = None

    # --- state -----------------------------------------------------------
    def set_state(self, cu, gp, time=0.0, total=0):
        cu = np.ascontiguousarray(cu, dtype=np.float64)
        gp = np.ascontiguousarray(gp, dtype=np.float64)
        lib().hyo_set_state(self.h, _dp(cu), _dp(gp), float(time), int(total))

    def get_state(self):
        cu = np.empty((10, self.ne))
        gp = np.empty((6, self.ng))
        t = ctypes.c_double()
        tot = ctypes.c_int()
        lib().hyo_get_state(self.h, _dp(cu), _dp(gp), ctypes.byref(t), ctypes.byref(tot))
        return cu, gp, t.value, tot.value

    def array(self, name):
        """numpy view of an internal array (shape (nvar, len))."""
        shapes = {"cn": (10, self.nt), "cu": (10, self.nt), "cn1": (10, self.nt),
                  "rhs": (10, self.ne), "d24": (10, self.ne), "d2q": (10, self.nt),
                  "anu": (2, self.nt), "q": (len(Q_NAMES), self.nt), "grad": (39, self.nt),
                  "phi": (self.nt,), "vis": (8, self.ne), "sc": (12, self.ne), "xc": (self.nt,),
                  "yc": (self.nt,), "zc": (self.nt,), "vol": (self.nt,), "dl": (self.ne,),
                  "dd": (self.nt,), "epssq": (self.ne,)}
        shp = shapes[name]
        ptr = lib().hyo_array(self.h, HA[name])
        n = int(np.prod(shp))
        buf = (ctypes.c_double * n).from_address(ptr)
        return np.frombuffer(buf, dtype=np.float64).reshape(shp)

    def q(self, name):
        return self.array("q")[Q_NAMES.index(name)]

    # --- kernels ---------------------------------------------------------
    def call(self, name, *args):
        return getattr(lib(), "hyo_" + name)(self.h, *args)

    def set_dt(self, dt, time, total):
        lib().hyo_set_dt(self.h, float(dt), float(time), int(total))

    @property
    def dt(self):
        return lib().hyo_get_dt(self.h)

    def step(self):
        d = np.zeros(4)
        lib().hyo_step(self.h, _dp(d))
        return d

    def run(self, n):
        lib().hyo_run(self.h, int(n))

    def face_states(self):
        """Reconstructed left/right face states before clipping, shape
        (neles, 4, 2, 17): face_state fields, then the face centroid."""
        out = np.empty((self.ne, 4, 2, 17))
        lib().hyo_face_states(self.h, _dp(out))
        return out

    def set_donors(self, recv, donors, weights):
        """Overset donor lists, 0-based (see hyo_set_donors)."""
        self._don = (np.ascontiguousarray(recv, dtype=np.int32), np.ascontiguousarray(donors, dtype=np.int32),
                     np.ascontiguousarray(weights, dtype=np.float64))
        r, d, w = self._don
        lib().hyo_set_donors(self.h, int(r.size), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(w))
