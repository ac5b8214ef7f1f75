"""ctypes binding of the CPU FP64 oracle (oracle/kitamr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libkitamr_oracle.so"
sys.path.insert(0, str(HERE.parent / "kitamr.jl_amd"))
from kitamr_hip import abi  # noqa: E402


def build(force=False):
    if force or not LIB.exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        dp = C.POINTER(C.c_double)
        L.orc_create.argtypes = [C.POINTER(abi.KahParams), C.POINTER(abi.KahTopology)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_state.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_get_state.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp]
        L.orc_get_flux.argtypes = [C.c_void_p, dp, dp]
        L.orc_slope.argtypes = [C.c_void_p]
        L.orc_flux.argtypes = [C.c_void_p, C.c_double]
        L.orc_iterate.argtypes = [C.c_void_p, C.c_double, C.c_int64, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_discrete_maxwell.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp, C.c_double, dp]
        L.orc_micro_to_macro.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp, dp, dp]
        L.orc_heat_flux.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp, dp, dp, dp]
        L.orc_i_projection.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp, dp, dp]
        L.orc_shakhov_part.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp, dp, dp,
                                       C.c_double, C.c_double, dp]
        L.orc_get_prim.argtypes = [C.c_int, dp, C.c_double, dp]
        L.orc_get_conserved.argtypes = [C.c_int, dp, C.c_double, dp]
        L.orc_get_tau.argtypes = [C.c_int, dp, C.c_double, C.c_double]
        L.orc_get_tau.restype = C.c_double
        L.orc_ref_vhs_vis.argtypes = [C.c_double] * 4
        L.orc_ref_vhs_vis.restype = C.c_double
        L.orc_vs_resolution.argtypes = [C.c_void_p, dp]
        L.orc_vs_flags.argtypes = [C.c_void_p, C.POINTER(abi.KahVsAdapt), dp,
                                   C.POINTER(C.c_uint8)]
        L.orc_set_sdf.argtypes = [C.c_void_p, dp]
        L.orc_get_ib.argtypes = [C.c_void_p, dp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def ib_sizes(topo):
    return abi.ib_sizes(topo)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The reference algorithm on the CPU, fed with the same topology as the device."""

    def __init__(self, params, topo):
        self.L = lib()
        s, keep = abi.make_topology_struct(topo)
        self.h = self.L.orc_create(C.byref(params), C.byref(s))
        del keep
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.dim, self.ndf = params.dim, params.ndf
        self.nc = int(topo["n_cells"])
        self.np = int(np.asarray(topo["vs_offset"])[-1])
        self.n_ib = int(len(topo.get("ib_face_pair", [])))
        self.n0 = int(np.asarray(topo["vs_offset"])[1]) if self.nc else 0
        self.ib_sizes = ib_sizes(topo)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {rc}: {self.L.orc_last_error(self.h).decode()}")

    def set_state(self, w, prim, df):
        self._ck(self.L.orc_set_state(self.h, _p(_f(w)), _p(_f(prim)), _p(_f(df))), "set_state")

    def get_state(self):
        nw = self.dim + 2
        a = dict(w=np.empty(self.nc * nw), prim=np.empty(self.nc * nw),
                 qf=np.empty(self.nc * self.dim), sw=np.empty(self.nc * nw * self.dim),
                 df=np.empty(self.np * self.ndf), sdf=np.empty(self.np * self.ndf * self.dim))
        self._ck(self.L.orc_get_state(self.h, *[_p(a[k]) for k in
                                                ["w", "prim", "qf", "sw", "df", "sdf"]]),
                 "get_state")
        return a

    def set_sdf(self, sdf):
        self._ck(self.L.orc_set_sdf(self.h, _p(_f(sdf))), "set_sdf")

    def vs_resolution(self):
        vr = np.empty(2)
        self._ck(self.L.orc_vs_resolution(self.h, _p(vr)), "vs_resolution")
        return vr

    def vs_flags(self, params, vr):
        out = np.empty(self.np, dtype=np.uint8)
        self._ck(self.L.orc_vs_flags(self.h, C.byref(params), _p(_f(vr)),
                                     out.ctypes.data_as(C.POINTER(C.c_uint8))), "vs_flags")
        return out

    def get_ib(self, flat=False):
        """SolidNeighbor distributions of the last flux: per face [ndf, n] on the donor's grid
        (an array [n_ib_faces, ndf, n0] for a shared grid); flat=True: the concatenation."""
        out = np.empty(max(int(self.ib_sizes.sum()) * self.ndf, 1))
        self._ck(self.L.orc_get_ib(self.h, _p(out)), "get_ib")
        return abi.split_ib(out[:int(self.ib_sizes.sum()) * self.ndf], self.ib_sizes, self.ndf, flat)

    def get_flux(self):
        f = np.empty(self.nc * (self.dim + 2))
        v = np.empty(self.np * self.ndf)
        self._ck(self.L.orc_get_flux(self.h, _p(f), _p(v)), "get_flux")
        return f, v

    def slope(self):
        self._ck(self.L.orc_slope(self.h), "slope")

    def flux(self, dt):
        self._ck(self.L.orc_flux(self.h, float(dt)), "flux")

    def iterate(self, dt, step):
        nw = self.dim + 2
        r, a = np.zeros(nw), np.zeros(nw)
        self._ck(self.L.orc_iterate(self.h, float(dt), int(step), _p(r), _p(a)), "iterate")
        return r, a

    def step(self, dt, step, residual=True):
        """slope!, flux!, iterate! of one step. The residual sums are always formed; with
        residual=False they are dropped, as the device's step does not form them then."""
        self.slope()
        self.flux(dt)
        r = self.iterate(dt, step)
        return r if residual else (None, None)


def set_threads(n):
    return lib().orc_set_threads(int(n))


# point kernels ------------------------------------------------------------------------
def discrete_maxwell(dim, ndf, mid_colmajor, prim, K):
    n = len(mid_colmajor) // dim
    out = np.empty(n * ndf)
    lib().orc_discrete_maxwell(dim, ndf, n, _p(_f(mid_colmajor)), _p(_f(prim)), K, _p(out))
    return out


def micro_to_macro(dim, ndf, mid_colmajor, df, weight):
    n = len(weight)
    out = np.empty(dim + 2)
    lib().orc_micro_to_macro(dim, ndf, n, _p(_f(mid_colmajor)), _p(_f(df)), _p(_f(weight)),
                             _p(out))
    return out


def heat_flux(dim, ndf, mid_colmajor, df, weight, prim):
    n = len(weight)
    out = np.empty(dim)
    lib().orc_heat_flux(dim, ndf, n, _p(_f(mid_colmajor)), _p(_f(df)), _p(_f(weight)),
                        _p(_f(prim)), _p(out))
    return out


def i_projection(dim, ndf, mid_colmajor, weight, df, w):
    """conserved_I_porjection! of one grid: returns the projected df ([ndf, n] flattened)."""
    n = len(weight)
    out = np.array(df, dtype=np.float64, copy=True)
    lib().orc_i_projection(dim, ndf, n, _p(_f(mid_colmajor)), _p(_f(weight)), _p(out),
                           _p(_f(w)))
    return out


def shakhov_part(dim, ndf, mid_colmajor, F, prim, qf, Pr, K):
    n = len(mid_colmajor) // dim
    out = np.empty(n * ndf)
    lib().orc_shakhov_part(dim, ndf, n, _p(_f(mid_colmajor)), _p(_f(F)), _p(_f(prim)),
                           _p(_f(qf)), Pr, K, _p(out))
    return out


def get_prim(dim, w, gamma):
    out = np.empty(dim + 2)
    lib().orc_get_prim(dim, _p(_f(w)), gamma, _p(out))
    return out


def get_conserved(dim, prim, gamma):
    out = np.empty(dim + 2)
    lib().orc_get_conserved(dim, _p(_f(prim)), gamma, _p(out))
    return out


def get_tau(dim, prim, mu, omega):
    return lib().orc_get_tau(dim, _p(_f(prim)), mu, omega)


def ref_vhs_vis(Kn, omega, alpha=1.0, T_ref=1.0):
    return lib().orc_ref_vhs_vis(Kn, omega, alpha, T_ref)
