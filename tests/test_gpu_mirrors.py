"""Ghost-exchange buffers on the device (kah_pack_mirrors / kah_unpack_mirrors) against the
wire format of src/Parallel/Ghost.jl.

The expected buffers are assembled here on the host from the reference's layout
(size_Ghost_Data / size_Ghost_Slope, :1-2; get_mirror_data, :525-557; get_mirror_slope_inner!,
:47-55; the ghost wrap that reads them back, :688-749) and the arrays kah_get_state
downloads: a transposed sw, a row-major df, a dropped vs_num or a shifted offset fails the
bit-for-bit comparison. The oracle's state packed the same way agrees to the step's parity
tolerance. Unpacking the buffers of a later state restores exactly the packed cells.
"""
import math

import numpy as np
import pytest

from _util import pair, rel_err, say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params

EPS = 1e-12   # src/Abstract/Types.jl:3

CASES = {
    "riemann8": lambda: cases.riemann_2d(n=8, nv=8),
    "riemann16_amr2": lambda: cases.riemann_2d(n=16, nv=12, ps_levels=2, halo=(3.0, 2.0)),
    "riemann8_vsamr": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2),
    "box3d_amr_vsamr": lambda: cases.box_3d(n=4, nv=6, ps_levels=1, vs_levels=1),
}


def host_pack(topo, st, cells, kind, dim, ndf):
    """Concatenated buffers of the listed cells in the reference's layout, from host arrays."""
    nw = dim + 2
    off = np.asarray(topo["vs_offset"])
    benc = np.asarray(topo["bound_enc"])
    ds = np.asarray(topo["cell_ds"]).reshape(-1, dim)
    mid = np.asarray(topo["cell_mid"]).reshape(-1, dim)
    out = []
    for c in cells:
        a, b = int(off[c]), int(off[c + 1])
        if kind == "data":
            if benc[c] < 0:
                out += [ds[c], mid[c], np.full(nw, np.inf), [0.0, float(benc[c])]]
            else:
                out += [ds[c], mid[c], st["w"].reshape(-1, nw)[c], [float(b - a), float(benc[c])],
                        st["df"][a * ndf:b * ndf]]
        else:
            if benc[c] < 0:
                out += [[EPS], np.zeros(dim * nw - 1)]
            else:
                sw = st["sw"].reshape(-1, nw, dim)[c]          # sw[q, d]
                out += [st["sdf"][a * ndf * dim:b * ndf * dim], sw.flatten(order="F")]
    return np.concatenate([np.asarray(x, dtype=np.float64) for x in out])


def _cells(nc, seed):
    rng = np.random.default_rng(seed)
    cells = rng.choice(nc, size=min(nc, 13), replace=False)
    return np.concatenate([cells, cells[:2]]).astype(np.int32)    # unsorted, with repeats


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_pack_matches_reference_layout(name, precision):
    case = CASES[name]()
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    say(f"[mirrors] pack {name} {precision}")
    dev, orc = pair(case, precision=abi.PREC[precision])
    dev.step(case.dt_xi, 0)
    orc.step(case.dt_xi, 0)
    cells = _cells(case.n_cells, 3)
    st = dev.get_state(sw=True, sdf=True)
    for kind in ("data", "slope"):
        got, off = dev.pack_mirrors(cells, kind)
        nv = np.diff(np.asarray(case.topo["vs_offset"]))[cells]
        size = 3 * dim + 4 + ndf * nv if kind == "data" else nv * ndf * dim + dim * (dim + 2)
        assert np.array_equal(np.diff(off), size)
        expect = host_pack(case.topo, st, cells, kind, dim, ndf)
        assert got.shape == expect.shape
        assert np.array_equal(got, expect), (kind, np.flatnonzero(got != expect)[:8])
    # the same buffers from the oracle's state: the step's parity tolerance (FP32 form: the
    # moment bound of the float storage)
    tol = 1e-12 if precision == "fp64" else 1e-5
    got, _ = dev.pack_mirrors(cells, "data")
    assert rel_err(got, host_pack(case.topo, orc.get_state(), cells, "data", dim, ndf)) < tol
    dev.close()
    orc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("name", ["riemann8", "riemann8_vsamr", "box3d_amr_vsamr"])
def test_unpack_restores_packed_cells(name, precision):
    case = CASES[name]()
    say(f"[mirrors] round trip {name} {precision}")
    params = kah_params(case.config)
    params.precision = abi.PREC[precision]
    dev = abi.HipContext(params)
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    dev.step(case.dt_xi, 0)
    cells = _cells(case.n_cells, 5)[:-2]
    data, _ = dev.pack_mirrors(cells, "data")
    slope, _ = dev.pack_mirrors(cells, "slope")
    before = dev.get_state(sdf=True)
    dev.step(case.dt_xi, 1)
    dev.step(case.dt_xi, 2)
    later = dev.get_state(sdf=True)
    dev.unpack_mirrors(cells, "data", data)
    dev.unpack_mirrors(cells, "slope", slope)
    assert np.array_equal(dev.pack_mirrors(cells, "data")[0], data)
    assert np.array_equal(dev.pack_mirrors(cells, "slope")[0], slope)
    now = dev.get_state(sdf=True)
    nw, ndf, dim = dev.dim + 2, dev.ndf, dev.dim
    off = np.asarray(case.topo["vs_offset"])
    listed = np.zeros(case.n_cells, bool)
    listed[cells] = True
    for c in range(case.n_cells):
        src = before if listed[c] else later
        a, b = int(off[c]), int(off[c + 1])
        assert np.array_equal(now["w"].reshape(-1, nw)[c], src["w"].reshape(-1, nw)[c])
        assert np.array_equal(now["df"][a * ndf:b * ndf], src["df"][a * ndf:b * ndf])
        assert np.array_equal(now["sdf"][a * ndf * dim:b * ndf * dim],
                              src["sdf"][a * ndf * dim:b * ndf * dim])
    dev.close()


@pytest.mark.gpu
def test_exchange_between_two_contexts():
    """Ghost update between two shards of one run: the receiving context's cells take the
    sender's state, and its next step equals the sender's step on those cells when every cell
    they read was exchanged (here: all of them)."""
    case = CASES["riemann16_amr2"]()
    say("[mirrors] two contexts")
    a, _ = pair(case)
    b, _ = pair(case)
    a.step(case.dt_xi, 0)
    a.step(case.dt_xi, 1)
    allc = np.arange(case.n_cells, dtype=np.int32)[::-1].copy()
    data, _ = a.pack_mirrors(allc, "data")
    b.unpack_mirrors(allc, "data", data)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["w"], sb["w"]) and np.array_equal(sa["df"], sb["df"])
    a.step(case.dt_xi, 2)
    b.step(case.dt_xi, 2)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["w"], sb["w"]) and np.array_equal(sa["df"], sb["df"])
    a.close()
    b.close()


@pytest.mark.gpu
def test_unpack_rejects_mismatched_buffers():
    case = CASES["riemann8_vsamr"]()
    say("[mirrors] rejects")
    dev, _ = pair(case)
    off = np.asarray(case.topo["vs_offset"])
    nv = np.diff(off)
    c0 = int(np.argmax(nv))
    c1 = int(np.argmin(nv))
    assert nv[c0] != nv[c1]
    data, _ = dev.pack_mirrors([c0], "data")
    before = dev.get_state()
    with pytest.raises(abi.KahError, match="vs_num"):
        dev.unpack_mirrors([c1], "data", np.concatenate([data, np.zeros(64)]))
    bad = data.copy()
    bad[2 * dev.dim] = math.inf
    with pytest.raises(abi.KahError, match="solid data for a fluid cell"):
        dev.unpack_mirrors([c0], "data", bad)
    with pytest.raises(abi.KahError, match="out of range"):
        dev.pack_mirrors([case.n_cells], "data")
    with pytest.raises(abi.KahError, match="unknown mirror buffer kind"):
        dev._check(dev.lib.kah_mirror_layout(dev.h, None, 0, 7, np.zeros(1, np.int64).ctypes
                                             .data_as(abi.C.POINTER(abi.C.c_int64))),
                   "kah_mirror_layout")
    after = dev.get_state()
    assert np.array_equal(before["df"], after["df"]) and np.array_equal(before["w"], after["w"])
    dev.close()


def _with_solid(case, c):
    """The case's topology with cell c inside a solid: faces owned by c are dropped (faces of
    its neighbours towards it stay, as solid faces)."""
    t = dict(case.topo)
    t["bound_enc"] = np.array(t["bound_enc"], copy=True)
    t["bound_enc"][c] = -1
    keep = np.asarray(t["pair_here"]) != c
    for k in ("pair_here", "pair_there", "pair_dir", "pair_rot", "pair_area", "pair_bc"):
        t[k] = np.asarray(t[k])[keep]
    t["pair_mid"] = np.asarray(t["pair_mid"]).reshape(-1, case.config.solver.DIM)[keep].ravel()
    t["n_pairs"] = int(keep.sum())
    return t


@pytest.mark.gpu
def test_solid_cells_pack_as_the_reference():
    case = cases.riemann_2d(n=8, nv=8)
    say("[mirrors] solid")
    topo = _with_solid(case, 27)
    dev = abi.HipContext(kah_params(case.config))
    dev.set_topology(topo)
    dev.set_state(case.w, case.prim, case.df)
    cells = np.array([26, 27, 28], np.int32)
    st = dev.get_state(sw=True, sdf=True)
    for kind in ("data", "slope"):
        got, off = dev.pack_mirrors(cells, kind)
        assert np.array_equal(got, host_pack(topo, st, cells, kind, 2, 2))
    d, off = dev.pack_mirrors(cells, "data")
    assert off[2] - off[1] == 3 * 2 + 4 and np.all(np.isinf(d[off[1] + 4:off[1] + 8]))
    # unpack skips the solid cell and still writes its neighbours
    before = dev.get_state()
    d2 = d.copy()
    d2[off[0] + 4] += 1.0                      # w[1] of cell 26
    dev.unpack_mirrors(cells, "data", d2)
    w = dev.get_state()["w"].reshape(-1, 4)
    assert w[26, 0] == before["w"].reshape(-1, 4)[26, 0] + 1.0
    assert np.array_equal(w[27], before["w"].reshape(-1, 4)[27])
    dev.close()
