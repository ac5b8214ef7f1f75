"""ctypes binding of the C ABI in include/kitamr_hip.h.

This is the same binding a Julia ``ccall`` shim performs: plain pointers and sizes, the
caller owning its host arrays. The library is loaded from ``kitamr.jl_amd/lib`` and the
load fails loudly when it is missing; nothing here falls back to a host implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libkitamr_hip.so"

FLUX = {"CAIDVM": 0, "DVM": 1}
MARCH = {"CAIDVM_Marching": 0, "Euler": 1, "CIP_Marching": 2}
PREC = {"fp64": 0, "fp32f": 1}
BC = {"Maxwellian": 0, "SuperSonicInflow": 1, "UniformOutflow": 2, "InterpolatedOutflow": 3}
ERRORS = {0: "ok", -1: "invalid argument", -2: "out of device memory", -3: "device error",
          -4: "bad call order", -5: "unsupported"}

EXPORTS = ["kah_create", "kah_destroy", "kah_last_error", "kah_set_topology", "kah_set_state",
           "kah_get_state", "kah_get_flux", "kah_slope", "kah_flux", "kah_iterate", "kah_step",
           "kah_sync", "kah_phase_cells", "kah_device_bytes", "kah_kernel_times",
           "kah_copy_bandwidth", "kah_set_option", "kah_mirror_layout", "kah_pack_mirrors",
           "kah_unpack_mirrors", "kah_restart_sizes", "kah_save_restart", "kah_load_restart",
           "kah_vs_resolution", "kah_vs_flags", "kah_vs_remap", "kah_vs_project", "kah_get_ib",
           "kah_vs_resolution_noalloc"]
MIRROR = {"data": 0, "slope": 1}
OPT = {"flux_waves": 1, "slope_vpt": 2, "vel_tiles": 3, "overlap_maps": 4, "graphs": 5,
       "regular": 6, "branches": 7}


class KahParams(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ndf", C.c_int32), ("flux_kind", C.c_int32),
                ("marching", C.c_int32), ("precision", C.c_int32),
                ("st_check_interval", C.c_int32), ("device", C.c_int32),
                ("reserved", C.c_int32), ("K", C.c_double), ("gamma", C.c_double),
                ("Pr", C.c_double), ("mu_ref", C.c_double), ("omega", C.c_double)]


class KahVsAdapt(C.Structure):
    _fields_ = [("mode", C.c_int32), ("reserved", C.c_int32), ("coeff_local", C.c_double),
                ("coeff_global", C.c_double), ("coeff_lohner", C.c_double)]


VS_MODE = {"lohner": 0, "contribution": 1}


def vs_adapt_params(mode="lohner", coeff_local=1e-2, coeff_global=0.125, coeff_lohner=0.6):
    """kah_vs_adapt with the reference's defaults (src/Solver/Types.jl:97-100)."""
    return KahVsAdapt(VS_MODE[mode], 0, coeff_local, coeff_global, coeff_lohner)


_P = C.c_void_p


class KahTopology(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_images", C.c_int64), ("image_src", _P),
                ("cell_mid", _P), ("cell_ds", _P), ("bound_enc", _P), ("ps_level", _P),
                ("vs_offset", _P), ("vs_level", _P), ("vs_mid", _P), ("vs_weight", _P),
                ("vs_uniform", C.c_int32), ("reserved0", C.c_int32),
                ("nb_state", _P), ("nb_offset", _P), ("nb_cell", _P),
                ("n_pairs", C.c_int64), ("pair_here", _P), ("pair_there", _P),
                ("pair_dir", _P), ("pair_rot", _P), ("pair_area", _P), ("pair_mid", _P),
                ("pair_bc", _P), ("n_bc", C.c_int32), ("vs_maxlevel", C.c_int32),
                ("bc_kind", _P), ("bc_prim", _P), ("vs_trees_num", C.c_int32 * 3),
                ("reserved", C.c_int32), ("vs_lo", C.c_double * 3), ("vs_hi", C.c_double * 3),
                ("n_ib_solid", C.c_int64), ("ib_solid", _P), ("ib_solid_nb_off", _P),
                ("ib_solid_nb", _P), ("n_ib_faces", C.c_int64), ("ib_face_pair", _P),
                ("ib_aux_point", _P), ("ib_normal", _P), ("ib_bc", _P), ("ib_image_off", _P),
                ("ib_image_cell", _P), ("ib_cut_off", _P), ("ib_cut_index", _P),
                ("ib_cut_gas", _P), ("ib_cut_solid", _P)]


_TOPO_ARRAYS = {
    "image_src": np.int64, "cell_mid": np.float64, "cell_ds": np.float64,
    "bound_enc": np.int32, "ps_level": np.int32, "vs_offset": np.int64, "vs_level": np.int8,
    "vs_mid": np.float64, "vs_weight": np.float64, "nb_state": np.int32,
    "nb_offset": np.int64, "nb_cell": np.int64, "pair_here": np.int64,
    "pair_there": np.int64, "pair_dir": np.int32, "pair_rot": np.float64,
    "pair_area": np.float64, "pair_mid": np.float64, "pair_bc": np.int32,
    "bc_kind": np.int32, "bc_prim": np.float64,
    "ib_solid": np.int64, "ib_solid_nb_off": np.int64, "ib_solid_nb": np.int64,
    "ib_face_pair": np.int64, "ib_aux_point": np.float64, "ib_normal": np.float64,
    "ib_bc": np.float64, "ib_image_off": np.int64, "ib_image_cell": np.int64,
    "ib_cut_off": np.int64, "ib_cut_index": np.int32, "ib_cut_gas": np.float64,
    "ib_cut_solid": np.float64,
}


def make_topology_struct(topo):
    """Fill a KahTopology from a dict of arrays; returns (struct, keepalive list)."""
    keep = []
    t = KahTopology()
    for name, dt in _TOPO_ARRAYS.items():
        a = topo.get(name)
        if a is None:
            setattr(t, name, None)
            continue
        a = np.ascontiguousarray(np.asarray(a, dtype=dt).reshape(-1))
        keep.append(a)
        setattr(t, name, a.ctypes.data if a.size else None)
    t.n_cells = int(topo["n_cells"])
    t.n_images = int(topo.get("n_images", 0))
    t.n_pairs = int(topo["n_pairs"])
    t.n_bc = int(len(topo.get("bc_kind", [])))
    t.vs_uniform = int(topo.get("vs_uniform", 0))
    t.vs_maxlevel = int(topo.get("vs_maxlevel", 0))
    t.n_ib_solid = int(len(topo.get("ib_solid", [])))
    t.n_ib_faces = int(len(topo.get("ib_face_pair", [])))
    vtn = list(topo.get("vs_trees_num", [0, 0, 0])) + [0, 0, 0]
    lo = list(topo.get("vs_lo", [0.0, 0.0, 0.0])) + [0.0, 0.0, 0.0]
    hi = list(topo.get("vs_hi", [0.0, 0.0, 0.0])) + [0.0, 0.0, 0.0]
    for d in range(3):
        t.vs_trees_num[d] = int(vtn[d])
        t.vs_lo[d] = float(lo[d])
        t.vs_hi[d] = float(hi[d])
    return t, keep


def ib_sizes(topo):
    """Velocity-grid size of each IB face's donor (the SolidNeighbor's grid)."""
    if not len(topo.get("ib_face_pair", [])):
        return np.zeros(0, np.int64)
    off = np.asarray(topo["vs_offset"])
    h = np.asarray(topo["pair_here"])[np.asarray(topo["ib_face_pair"])]
    return (off[h + 1] - off[h]).astype(np.int64)


def split_ib(flat_arr, sizes, ndf, flat=False):
    """Per-face [ndf, n] blocks of the concatenated SolidNeighbor distributions."""
    if flat:
        return flat_arr
    if len(sizes) and np.all(sizes == sizes[0]):
        return flat_arr.reshape(len(sizes), ndf, int(sizes[0]))
    out, o = [], 0
    for n in sizes:
        out.append(flat_arr[o:o + ndf * n].reshape(ndf, n))
        o += ndf * n
    return out


def make_params(dim, ndf, flux="CAIDVM", marching="CAIDVM_Marching", precision="fp64",
                K=1.0, gamma=5.0 / 3.0, Pr=2.0 / 3.0, mu_ref=0.0, omega=0.5,
                st_check_interval=100, device=0):
    p = KahParams()
    p.dim, p.ndf = dim, ndf
    p.flux_kind = FLUX[flux]
    p.marching = MARCH[marching]
    p.precision = PREC[precision]
    p.st_check_interval = st_check_interval
    p.device = device
    p.K, p.gamma, p.Pr, p.mu_ref, p.omega = K, gamma, Pr, mu_ref, omega
    return p


_lib = None


def load_library(path=None):
    """Load libkitamr_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    d, i32, i64 = C.c_double, C.c_int32, C.c_int64
    dp = C.POINTER(C.c_double)
    lib.kah_create.argtypes = [C.POINTER(KahParams), C.POINTER(C.c_void_p)]
    lib.kah_create.restype = C.c_int
    lib.kah_destroy.argtypes = [C.c_void_p]
    lib.kah_destroy.restype = None
    lib.kah_last_error.argtypes = [C.c_void_p]
    lib.kah_last_error.restype = C.c_char_p
    lib.kah_set_topology.argtypes = [C.c_void_p, C.POINTER(KahTopology)]
    lib.kah_set_state.argtypes = [C.c_void_p, dp, dp, dp]
    lib.kah_get_state.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp]
    lib.kah_get_flux.argtypes = [C.c_void_p, dp, dp]
    lib.kah_get_ib.argtypes = [C.c_void_p, dp]
    lib.kah_slope.argtypes = [C.c_void_p, d]
    lib.kah_flux.argtypes = [C.c_void_p, d]
    lib.kah_iterate.argtypes = [C.c_void_p, d, i64, dp, dp]
    lib.kah_step.argtypes = [C.c_void_p, d, i64, dp, dp]
    lib.kah_sync.argtypes = [C.c_void_p]
    lib.kah_phase_cells.argtypes = [C.c_void_p]
    lib.kah_phase_cells.restype = i64
    lib.kah_device_bytes.argtypes = [C.c_void_p]
    lib.kah_device_bytes.restype = i64
    lib.kah_kernel_times.argtypes = [C.c_void_p, C.c_int, dp, C.POINTER(i64)]
    lib.kah_copy_bandwidth.argtypes = [C.c_void_p, i64, C.c_int, dp]
    lib.kah_copy_bandwidth.restype = C.c_int
    lib.kah_set_option.argtypes = [C.c_void_p, C.c_int, i64]
    lib.kah_set_option.restype = C.c_int
    ip = C.POINTER(C.c_int32)
    lib.kah_mirror_layout.argtypes = [C.c_void_p, ip, i64, C.c_int, C.POINTER(i64)]
    lib.kah_pack_mirrors.argtypes = [C.c_void_p, ip, i64, C.c_int, dp]
    lib.kah_unpack_mirrors.argtypes = [C.c_void_p, ip, i64, C.c_int, dp]
    lp = C.POINTER(i64)
    lib.kah_restart_sizes.argtypes = [C.c_void_p, lp, lp]
    lib.kah_save_restart.argtypes = [C.c_void_p, lp, lp, dp, C.POINTER(C.c_int8), dp, dp]
    lib.kah_load_restart.argtypes = [C.c_void_p, lp, lp, dp, dp]
    u8p = C.POINTER(C.c_uint8)
    lib.kah_vs_resolution.argtypes = [C.c_void_p, dp]
    lib.kah_vs_resolution_noalloc.argtypes = [C.c_void_p, dp]
    lib.kah_vs_flags.argtypes = [C.c_void_p, C.POINTER(KahVsAdapt), dp, u8p]
    lib.kah_vs_remap.argtypes = [C.c_void_p, C.POINTER(KahTopology), ip, u8p]
    lib.kah_vs_project.argtypes = [C.c_void_p, ip, i64]
    for name in ["kah_set_topology", "kah_set_state", "kah_get_state", "kah_get_flux",
                 "kah_slope", "kah_flux", "kah_iterate", "kah_step", "kah_sync",
                 "kah_kernel_times", "kah_mirror_layout", "kah_pack_mirrors",
                 "kah_unpack_mirrors", "kah_restart_sizes", "kah_save_restart",
                 "kah_load_restart", "kah_vs_resolution", "kah_vs_flags", "kah_vs_remap",
                 "kah_vs_project", "kah_get_ib", "kah_vs_resolution_noalloc"]:
        getattr(lib, name).restype = C.c_int
    del i32
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class KahError(RuntimeError):
    pass


class HipContext:
    """One device context: owns the device state of one flow case."""

    def __init__(self, params, lib=None):
        self.lib = lib or load_library()
        self.params = params
        self.dim, self.ndf = params.dim, params.ndf
        h = C.c_void_p()
        rc = self.lib.kah_create(C.byref(params), C.byref(h))
        if rc != 0:
            raise KahError(f"kah_create failed: {ERRORS.get(rc, rc)}")
        self.h = h
        self.nc = 0
        self.np = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.kah_last_error(self.h).decode()
            raise KahError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.kah_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, topo):
        s, keep = make_topology_struct(topo)
        self._check(self.lib.kah_set_topology(self.h, C.byref(s)), "kah_set_topology")
        del keep
        self.nc = int(topo["n_cells"])
        off = np.asarray(topo["vs_offset"])
        self.np = int(off[-1])
        self.n_ib = int(len(topo.get("ib_face_pair", [])))
        self.n0 = int(off[1]) if self.nc else 0
        self.ib_sizes = ib_sizes(topo)

    def set_state(self, w, prim, df):
        w = np.ascontiguousarray(w, dtype=np.float64)
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        df = np.ascontiguousarray(df, dtype=np.float64)
        self._check(self.lib.kah_set_state(self.h, _ptr(w), _ptr(prim), _ptr(df)), "kah_set_state")

    def get_state(self, w=True, prim=True, qf=True, sw=False, df=True, sdf=False):
        nw = self.dim + 2
        out = {}
        arrs = {
            "w": np.empty(self.nc * nw) if w else None,
            "prim": np.empty(self.nc * nw) if prim else None,
            "qf": np.empty(self.nc * self.dim) if qf else None,
            "sw": np.empty(self.nc * nw * self.dim) if sw else None,
            "df": np.empty(self.np * self.ndf) if df else None,
            "sdf": np.empty(self.np * self.ndf * self.dim) if sdf else None,
        }
        self._check(self.lib.kah_get_state(self.h, *[_ptr(arrs[k]) for k in
                                                     ["w", "prim", "qf", "sw", "df", "sdf"]]),
                    "kah_get_state")
        for k, v in arrs.items():
            if v is not None:
                out[k] = v
        return out

    def get_flux(self):
        flux = np.empty(self.nc * (self.dim + 2))
        vflux = np.empty(self.np * self.ndf)
        self._check(self.lib.kah_get_flux(self.h, _ptr(flux), _ptr(vflux)), "kah_get_flux")
        return flux, vflux

    def get_ib(self, flat=False):
        """The SolidNeighbor distributions of the last flux: per face [ndf, n] on the donor's
        grid (an array [n_ib_faces, ndf, n0] for a shared grid); flat=True: the
        concatenation."""
        sizes = getattr(self, "ib_sizes", np.zeros(0, np.int64))
        tot = int(sizes.sum()) * self.ndf
        out = np.empty(max(tot, 1))
        self._check(self.lib.kah_get_ib(self.h, _ptr(out)), "kah_get_ib")
        return split_ib(out[:tot], sizes, self.ndf, flat)

    def _cells(self, cells):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        return c, c.ctypes.data_as(C.POINTER(C.c_int32))

    def mirror_layout(self, cells, kind):
        """Offsets (n+1, in doubles) of the listed cells' ghost-exchange buffers."""
        c, cp = self._cells(cells)
        off = np.empty(len(c) + 1, dtype=np.int64)
        self._check(self.lib.kah_mirror_layout(self.h, cp, len(c), MIRROR[kind],
                                               off.ctypes.data_as(C.POINTER(C.c_int64))),
                    "kah_mirror_layout")
        return off

    def pack_mirrors(self, cells, kind):
        """Concatenated DATA or SLOPE buffers of the listed cells (get_mirror_data /
        get_mirror_slope, src/Parallel/Ghost.jl:525-579) and their offsets."""
        off = self.mirror_layout(cells, kind)
        c, cp = self._cells(cells)
        out = np.empty(int(off[-1]))
        self._check(self.lib.kah_pack_mirrors(self.h, cp, len(c), MIRROR[kind], _ptr(out)),
                    "kah_pack_mirrors")
        return out, off

    def unpack_mirrors(self, cells, kind, buf):
        """Writes the listed cells' state from concatenated DATA or SLOPE buffers."""
        c, cp = self._cells(cells)
        b = np.ascontiguousarray(buf, dtype=np.float64)
        self._check(self.lib.kah_unpack_mirrors(self.h, cp, len(c), MIRROR[kind], _ptr(b)),
                    "kah_unpack_mirrors")

    def save_restart(self):
        """Per-rank restart arrays (save_for_restart, src/IO/Restart.jl:80-117) as numpy arrays
        of the reference's logical shapes: ws (N, dim+2), vs_midpoints (Nv, dim),
        vs_df (Nv, ndf)."""
        n, nv = C.c_int64(), C.c_int64()
        self._check(self.lib.kah_restart_sizes(self.h, C.byref(n), C.byref(nv)),
                    "kah_restart_sizes")
        N, Nv, nw = n.value, nv.value, self.dim + 2
        a = dict(vs_nums=np.empty(N, np.int64), bound_encs=np.empty(N, np.int64),
                 ws=np.empty(N * nw), vs_levels=np.empty(Nv, np.int8),
                 vs_midpoints=np.empty(Nv * self.dim), vs_df=np.empty(Nv * self.ndf))
        lp = lambda x: x.ctypes.data_as(C.POINTER(C.c_int64))
        self._check(self.lib.kah_save_restart(
            self.h, lp(a["vs_nums"]), lp(a["bound_encs"]), _ptr(a["ws"]),
            a["vs_levels"].ctypes.data_as(C.POINTER(C.c_int8)), _ptr(a["vs_midpoints"]),
            _ptr(a["vs_df"])), "kah_save_restart")
        # column-major matrices -> numpy arrays of the same logical shape
        a["ws"] = a["ws"].reshape(nw, N).T
        a["vs_midpoints"] = a["vs_midpoints"].reshape(self.dim, Nv).T
        a["vs_df"] = a["vs_df"].reshape(self.ndf, Nv).T
        return a

    def load_restart(self, arrays):
        """Sets the state from restart arrays of this topology (logical shapes as returned by
        save_restart)."""
        vn = np.ascontiguousarray(arrays["vs_nums"], dtype=np.int64)
        be = np.ascontiguousarray(arrays["bound_encs"], dtype=np.int64)
        ws = np.ascontiguousarray(np.asarray(arrays["ws"], dtype=np.float64).T)
        df = np.ascontiguousarray(np.asarray(arrays["vs_df"], dtype=np.float64).T)
        lp = lambda x: x.ctypes.data_as(C.POINTER(C.c_int64))
        self._check(self.lib.kah_load_restart(self.h, lp(vn), lp(be), _ptr(ws), _ptr(df)),
                    "kah_load_restart")

    def vs_resolution(self):
        """Velocity_Resolution (density, energy) of the current state."""
        vr = np.empty(2)
        self._check(self.lib.kah_vs_resolution(self.h, _ptr(vr)), "kah_vs_resolution")
        return vr

    def vs_resolution_noalloc(self):
        """vs_resolution through the allocation-free entry point (fixed device scratch)."""
        vr = np.empty(2)
        self._check(self.lib.kah_vs_resolution_noalloc(self.h, _ptr(vr)),
                    "kah_vs_resolution_noalloc")
        return vr

    def vs_flags(self, params, vr):
        """Per-velocity-cell flag bytes (1: refine, 2: coarsen) of the current grids."""
        out = np.empty(self.np, dtype=np.uint8)
        v = np.ascontiguousarray(vr, dtype=np.float64)
        self._check(self.lib.kah_vs_flags(self.h, C.byref(params), _ptr(v),
                                          out.ctypes.data_as(C.POINTER(C.c_uint8))),
                    "kah_vs_flags")
        return out

    def vs_remap(self, topo, src, changed):
        """Replaces the velocity grids by those of topo, carrying the state over."""
        s, keep = make_topology_struct(topo)
        sv = np.ascontiguousarray(src, dtype=np.int32)
        ch = np.ascontiguousarray(changed, dtype=np.uint8)
        self._check(self.lib.kah_vs_remap(self.h, C.byref(s), sv.ctypes.data_as(C.POINTER(C.c_int32)),
                                          ch.ctypes.data_as(C.POINTER(C.c_uint8))),
                    "kah_vs_remap")
        del keep
        self.nc = int(topo["n_cells"])
        self.np = int(np.asarray(topo["vs_offset"])[-1])
        self.ib_sizes = ib_sizes(topo)   # the SolidNeighbors follow their donors' new grids

    def vs_project(self, cells):
        c, cp = self._cells(cells)
        self._check(self.lib.kah_vs_project(self.h, cp, len(c)), "kah_vs_project")

    def slope(self, dt):
        self._check(self.lib.kah_slope(self.h, float(dt)), "kah_slope")

    def flux(self, dt):
        self._check(self.lib.kah_flux(self.h, float(dt)), "kah_flux")

    def iterate(self, dt, step):
        nw = self.dim + 2
        r, a = np.zeros(nw), np.zeros(nw)
        self._check(self.lib.kah_iterate(self.h, float(dt), int(step), _ptr(r), _ptr(a)),
                    "kah_iterate")
        return r, a

    def step(self, dt, step, residual=True):
        nw = self.dim + 2
        r, a = (np.zeros(nw), np.zeros(nw)) if residual else (None, None)
        self._check(self.lib.kah_step(self.h, float(dt), int(step), _ptr(r), _ptr(a)), "kah_step")
        return r, a

    def sync(self):
        self._check(self.lib.kah_sync(self.h), "kah_sync")

    def phase_cells(self):
        return int(self.lib.kah_phase_cells(self.h))

    def device_bytes(self):
        return int(self.lib.kah_device_bytes(self.h))

    def set_option(self, name, value):
        self._check(self.lib.kah_set_option(self.h, OPT[name], int(value)), "kah_set_option")

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        """Attainable HBM bandwidth (GB/s) of a plain streaming copy of nbytes."""
        ms = np.zeros(1)
        self._check(self.lib.kah_copy_bandwidth(self.h, int(nbytes), int(reps), _ptr(ms)),
                    "kah_copy_bandwidth")
        return 2.0 * nbytes / (ms[0] * 1e-3) / 1e9, float(ms[0])

    def kernel_times(self, reset=False):
        ms = np.zeros(2)
        calls = np.zeros(2, dtype=np.int64)
        self._check(self.lib.kah_kernel_times(self.h, int(reset), _ptr(ms),
                                              calls.ctypes.data_as(C.POINTER(C.c_int64))),
                    "kah_kernel_times")
        return ms, calls


def library_exports(path=None):
    """Names of the kah_* entry points the shared object exports (no device call)."""
    p = Path(path) if path else LIB_PATH
    lib = C.CDLL(str(p))
    return [n for n in EXPORTS if hasattr(lib, n)]


os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")
