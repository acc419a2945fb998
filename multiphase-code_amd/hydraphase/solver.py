"""Host-side mirror of the reference's per-step call surface.

The reference drives argument-less subroutines over module globals
(main.F90:511-657): primitive, the dt loop, SMOOTH, RK (-> VISCOUSGP,
CallFluxSolver -> hllcbatten/hllczien), PRIMITIVE.  A Solver owns one
device-resident case and exposes the same routines by name; every call
goes to libhydra_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import HydraError

ARRAY_ID = {"cn": 0, "cu": 1, "cn1": 2, "rhs": 3, "d24": 4, "d2q": 5, "anu": 6, "q": 7, "grad": 8,
            "phi": 9, "vis": 10, "dl": 12, "grad32": 13, "phi32": 14}
Q_NAMES = ["ug", "vg", "wg", "pg", "pp", "tg", "phig", "up", "vp", "wp", "phip", "tp", "rop", "rog"]
G_NAMES = ["ug", "vg", "wg", "tg", "up", "vp", "wp", "tp", "pg", "pp", "phig"]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Solver:
    def __init__(self, mesh, params: dict):
        L = _lib.lib()
        self.L = L
        self.mesh = mesh
        self.params = dict(params)
        ne, ng = mesh.neles, mesh.nghosts
        self.ne, self.ng, self.nt = ne, ng, ne + ng
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        nod = i32(mesh.nod + 1)
        nc = i32(mesh.nc + 1)
        self._keep = keep = dict(
            x=np.ascontiguousarray(mesh.x, dtype=np.float64), y=np.ascontiguousarray(mesh.y, dtype=np.float64),
            z=np.ascontiguousarray(mesh.z, dtype=np.float64), nod=nod, nc=nc, ib=i32(mesh.iblank),
            gpar=i32(mesh.gparent + 1), gid=i32(np.arange(ng) + ne + 1), gtype=i32(mesh.gtype),
            gside=i32(mesh.gside))
        self._p = _lib.make_params(params)
        h = ctypes.c_void_p()
        rc = L.hp_create(ctypes.byref(h), mesh.nodes, ne, ng, _dp(keep["x"]), _dp(keep["y"]), _dp(keep["z"]),
                         _ip(nod), ne, _ip(nc[0]), _ip(nc[1]), _ip(nc[2]), _ip(nc[3]), _ip(keep["ib"]),
                         _ip(keep["gpar"]), _ip(keep["gid"]), _ip(keep["gtype"]), _ip(keep["gside"]),
                         ctypes.byref(self._p))
        if rc != 0:
            raise HydraError(f"hp_create failed ({rc}): {L.hp_last_error(None).decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.hp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise HydraError(f"hydra call failed ({rc}): {self.L.hp_last_error(self.h).decode()}")

    # -- state (restart%04d.in) ---------------------------------------------
    def upload(self, cu, gp, time=0.0, total=0):
        cu = np.ascontiguousarray(cu, dtype=np.float64)
        gp = np.ascontiguousarray(gp, dtype=np.float64)
        assert cu.shape == (10, self.ne) and gp.shape == (6, self.ng)
        self._chk(self.L.hp_upload_state(self.h, _dp(cu), self.ne, _dp(gp), float(time), int(total)))

    def download(self):
        cu = np.empty((10, self.ne))
        gp = np.empty((6, self.ng))
        t = ctypes.c_double()
        tot = ctypes.c_int()
        self._chk(self.L.hp_download_state(self.h, _dp(cu), self.ne, _dp(gp), ctypes.byref(t), ctypes.byref(tot)))
        return cu, gp, t.value, tot.value

    # -- the reference's routines ------------------------------------------
    def primitive(self):
        self._chk(self.L.hp_primitive(self.h))

    def dt_min(self):
        v = ctypes.c_double()
        self._chk(self.L.hp_dt_min(self.h, ctypes.byref(v)))
        return v.value

    def set_dt(self, dt, time, total):
        self._chk(self.L.hp_set_dt(self.h, float(dt), float(time), int(total)))

    def smooth(self):
        self._chk(self.L.hp_smooth(self.h))

    def viscousgp(self):
        self._chk(self.L.hp_viscous(self.h))

    def flux(self):
        """CallFluxSolver: primitiveplusbc, derivative, limiter, face loop,
        source; leaves the residual in rhs."""
        self._chk(self.L.hp_flux(self.h))

    def rk_stage(self, s):
        self._chk(self.L.hp_rk_stage(self.h, int(s)))

    def rk(self):
        self._chk(self.L.hp_rk(self.h))

    def end_step(self, with_diffs=True):
        d = np.zeros(4)
        self._chk(self.L.hp_end_step(self.h, _dp(d) if with_diffs else None))
        return d

    def step(self, n=1):
        self._chk(self.L.hp_step(self.h, int(n)))

    def step_diffs(self, n=1):
        """n fused steps; returns each step's monitor (n, 4): diffu, diffv,
        diffw, diffp (main.F90:653)."""
        d = np.zeros((int(n), 4))
        self._chk(self.L.hp_step_diffs(self.h, int(n), _dp(d)))
        return d

    def last_diffs(self):
        d = np.zeros(4)
        self._chk(self.L.hp_last_diffs(self.h, _dp(d)))
        return d

    def set_flux_timing(self, on=True):
        self._chk(self.L.hp_set_flux_timing(self.h, 1 if on else 0))

    def synchronize(self):
        self._chk(self.L.hp_synchronize(self.h))

    def health(self) -> int:
        """0 if the state is healthy, else HEALTH_NONFINITE | HEALTH_PRESSURE
        bits (hp_health)."""
        v = ctypes.c_int()
        self._chk(self.L.hp_health(self.h, ctypes.byref(v)))
        return v.value

    # -- overset -----------------------------------------------------------
    def set_donors(self, recv, donors, weights):
        """Fixed donor lists (0-based receptor ids (n,), donors (n, 4),
        weights (n, 4)); applied after every RK stage update."""
        recv = np.ascontiguousarray(np.asarray(recv) + 1, dtype=np.int32)
        n = recv.size
        d = np.ascontiguousarray((np.asarray(donors) + 1).T, dtype=np.int32)   # (4, n) = column-major (n, 4)
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).T)
        self._donors = (recv, d, w)
        self._chk(self.L.hp_set_donors(self.h, n, _ip(recv), _ip(d), n, _dp(w)))

    def apply_donors(self):
        self._chk(self.L.hp_apply_donors(self.h))

    def last_timing(self):
        t = ctypes.c_double()
        f = ctypes.c_double()
        n = ctypes.c_int()
        self._chk(self.L.hp_last_timing(self.h, ctypes.byref(t), ctypes.byref(f), ctypes.byref(n)))
        return t.value, f.value, n.value

    # -- raw arrays --------------------------------------------------------
    def _shape(self, name):
        ne, nt = self.ne, self.nt
        return {"cn": (10, ne), "cu": (10, ne), "cn1": (10, ne), "rhs": (10, ne), "d24": (10, ne),
                "d2q": (10, ne), "anu": (2, ne), "q": (14, nt), "grad": (33, ne), "phi": (ne,),
                "vis": (8, ne), "dl": (ne,), "grad32": (33, ne), "phi32": (ne,)}[name]

    def get(self, name):
        if name == "scalars":
            out = np.zeros(3)
            self._chk(self.L.hp_get_array(self.h, 11, _dp(out), 3))
            return out
        out = np.empty(self._shape(name))
        self._chk(self.L.hp_get_array(self.h, ARRAY_ID[name], _dp(out), out.size))
        return out

    def set(self, name, value):
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(self._shape(name))
        self._chk(self.L.hp_set_array(self.h, ARRAY_ID[name], _dp(v), v.size))
