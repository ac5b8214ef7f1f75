"""Parity of the HIP path with the oracle (FP64): slope, flux and iterate separately, the
fused step, and several steps, on small cases of every supported configuration."""
import numpy as np
import pytest

from _util import pair, rel_err, rel_err_cols, rel_err_rows, say
from kitamr_hip import cases

TOL = 1e-12   # relative, FP64 form (scale-aware, per array)

CASES = {
    "riemann8": lambda: cases.riemann_2d(n=8, nv=8),
    "riemann16_amr2": lambda: cases.riemann_2d(n=16, nv=12, ps_levels=2, halo=(3.0, 2.0)),
    "wall_channel": lambda: cases.wall_channel(n=8, nv=8),
    "wall_channel_amr": lambda: cases.wall_channel(n=8, nv=8, ps_levels=1),
    "riemann8_dvm": lambda: cases.riemann_2d(n=8, nv=8, flux="DVM"),
    "riemann8_euler": lambda: cases.riemann_2d(n=8, nv=8, marching="Euler"),
    "riemann8_inflow": lambda: cases.riemann_2d(n=8, nv=8, bc="SuperSonicInflow"),
    "riemann8_interp": lambda: cases.riemann_2d(n=8, nv=8, bc="InterpolatedOutflow"),
    "riemann8_wall": lambda: cases.riemann_2d(n=8, nv=8, bc="Maxwellian"),
    "riemann8_vsamr": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2),
    "riemann16_amr_vsamr": lambda: cases.riemann_2d(n=8, nv=8, ps_levels=2, halo=(2.0, 1.5),
                                                     vs_levels=1),
    "box3d": lambda: cases.box_3d(n=4, nv=6),
    "box3d_amr_vsamr": lambda: cases.box_3d(n=4, nv=6, ps_levels=1, vs_levels=1),
    # CIP_Marching (conserved I-projection)
    "riemann8_cip": lambda: cases.riemann_2d(n=8, nv=16, marching="CIP_Marching"),
    "riemann16_amr2_cip": lambda: cases.riemann_2d(n=16, nv=12, ps_levels=2, halo=(3.0, 2.0),
                                                   marching="CIP_Marching"),
    "riemann8_vsamr_cip": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2,
                                                   marching="CIP_Marching"),
    "box3d_cip": lambda: cases.box_3d(n=4, nv=12, marching="CIP_Marching"),
    # grids whose size admits wave tiles (8x8 in 2D, 4x4x4 in 3D)
    "riemann8_amr1_nv16": lambda: cases.riemann_2d(n=8, nv=16, ps_levels=1, halo=(2.0, 1.5)),
    "box3d_nv8": lambda: cases.box_3d(n=4, nv=8),
}

# launch options that change the schedule but not the arithmetic of each update
OPTION_SETS = {
    "no_tiles": {"vel_tiles": 0},
    "vpt1": {"slope_vpt": 1},
    "vpt3": {"slope_vpt": 3},
    "vpt4": {"slope_vpt": 4},
    "vpt5": {"slope_vpt": 5},
    "vpt6": {"slope_vpt": 6},
    "waves4": {"flux_waves": 4},
    "waves3": {"flux_waves": 3},
    "waves2": {"flux_waves": 2},
    "no_maps": {"overlap_maps": 0},
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_split_phases(name):
    case = CASES[name]()
    say(f"[split] {name}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = pair(case)
    dt = case.dt_xi
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    dev.slope(dt)
    orc.slope()
    g, o = dev.get_state(sdf=True, sw=True), orc.get_state()
    assert rel_err(g["sdf"], o["sdf"]) <= TOL, "sdf"
    assert rel_err_rows(g["sw"], o["sw"], (dim + 2) * dim) <= 1e-10, "sw"
    dev.flux(dt)
    orc.flux(dt)
    gf, gv = dev.get_flux()
    of, ov = orc.get_flux()
    assert rel_err(gv, ov) <= TOL, "vs flux"
    assert rel_err(gf, of) <= TOL, "macro flux"
    rg = dev.iterate(dt, 0)
    ro = orc.iterate(dt, 0)
    g, o = dev.get_state(), orc.get_state()
    assert rel_err(g["df"], o["df"]) <= TOL, "df"
    assert rel_err_rows(g["w"], o["w"], dim + 2) <= TOL, "w"
    assert rel_err_rows(g["prim"], o["prim"], dim + 2) <= TOL, "prim"
    assert rel_err(g["qf"], o["qf"]) <= 1e-10, "qf"
    assert rel_err(rg[1], ro[1]) <= TOL
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_fused_steps(name):
    case = CASES[name]()
    say(f"[fused] {name}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = pair(case)
    dt = case.dt_xi
    dim = case.config.solver.DIM
    for s in range(4):
        rg = dev.step(dt, s)
        ro = orc.step(dt, s)
    g, o = dev.get_state(), orc.get_state()
    assert rel_err(g["df"], o["df"]) <= TOL, "df"
    assert rel_err_rows(g["w"], o["w"], dim + 2) <= TOL, "w"
    assert rel_err_rows(g["prim"], o["prim"], dim + 2) <= TOL, "prim"
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("opt", list(OPTION_SETS))
@pytest.mark.parametrize("name", ["riemann8_amr1_nv16", "box3d_nv8", "riemann16_amr2",
                                  "riemann16_amr_vsamr", "box3d_amr_vsamr"])
def test_launch_options(name, opt):
    case = CASES[name]()
    say(f"[options] {name} {opt}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = pair(case, options=OPTION_SETS[opt])
    dt = case.dt_xi
    dim = case.config.solver.DIM
    for s in range(3):
        dev.step(dt, s)
        orc.step(dt, s)
    g, o = dev.get_state(), orc.get_state()
    assert rel_err(g["df"], o["df"]) <= TOL, "df"
    assert rel_err_rows(g["w"], o["w"], dim + 2) <= TOL, "w"
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["riemann8_vsamr", "riemann16_amr_vsamr", "box3d_amr_vsamr"])
def test_overlap_maps_equal_search(name):
    """Adapted grids: the precomputed overlap maps give bit-identical steps to the in-kernel
    key search (same overlapping cells, factors and order; same thread order)."""
    case = CASES[name]()
    say(f"[maps] {name}: cells={case.n_cells} P={case.phase_cells}")
    a, _ = pair(case)
    b, _ = pair(case, options={"overlap_maps": 0})
    for s in range(3):
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
    ga, gb = a.get_state(sdf=True), b.get_state(sdf=True)
    for k in ("df", "sdf", "w", "prim"):
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["riemann8_amr1_nv16", "riemann16_amr_vsamr", "riemann8_cip",
                                  "box3d_nv8", "wall_channel"])
def test_graph_replay_equals_launches(name):
    """kah_step replayed from captured graphs (one per buffer parity and dt) gives the same
    state as the launches one by one, step after step, with host reads between the replays
    and a change of dt (recapture)."""
    case = CASES[name]()
    say(f"[graph] {name}: cells={case.n_cells} P={case.phase_cells}")
    a, _ = pair(case)
    b, _ = pair(case, options={"graphs": 1})
    dts = [case.dt_xi] * 4 + [0.5 * case.dt_xi] * 2 + [case.dt_xi] * 2
    for s, dt in enumerate(dts):
        ra = a.step(dt, s)
        rb = b.step(dt, s)
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
        ga, gb = a.get_state(), b.get_state()
        for k in ("df", "w", "prim", "qf"):
            assert np.array_equal(ga[k], gb[k]), (s, k)
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["riemann8", "riemann8_euler", "riemann8_dvm", "riemann8_cip",
                                  "riemann16_amr2", "box3d", "box3d_cip", "cylinder_c3_32",
                                  "riemann8_vsamr", "box3d_amr_vsamr"])
@pytest.mark.parametrize("precision", [0, 1])
def test_regular_cells_equal_general_gather(name, precision):
    """The specialised slope and gather of regular cells (one neighbour per side, faces to
    fluid cells on the same level, on the same or another velocity grid) give the same steps as
    the general ones, bit for bit, in both forms."""
    if name == "cylinder_c3_32":
        case = cases.cylinder_c3(n=32)
    else:
        case = CASES[name]()
    say(f"[regular] {name} precision {precision}: cells={case.n_cells} P={case.phase_cells}")
    a, _ = pair(case, options={"regular": 0}, precision=precision)
    b, _ = pair(case, options={"regular": 1}, precision=precision)
    for s in range(3):
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
    ga, gb = a.get_state(sdf=True), b.get_state(sdf=True)
    for k in ("df", "sdf", "w", "prim", "qf"):
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["riemann16_amr2", "riemann16_amr_vsamr", "box3d_amr_vsamr",
                                  "cylinder_c3_32"])
@pytest.mark.parametrize("graphs", [0, 1])
def test_branches_equal_one_stream(name, graphs):
    """The independent launches of a phase spread over four streams (forked and joined by
    events; parallel branches of the captured graph) give the same state as one stream, bit
    for bit, step after step: every launch writes its own cells."""
    case = cases.cylinder_c3(n=32) if name == "cylinder_c3_32" else CASES[name]()
    say(f"[branches] {name} graphs={graphs}: cells={case.n_cells} P={case.phase_cells}")
    a, _ = pair(case, options={"graphs": graphs})
    b, _ = pair(case, options={"graphs": graphs, "branches": 4})
    for s in range(4):
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
        ga, gb = a.get_state(sdf=True), b.get_state(sdf=True)
        for k in ("df", "sdf", "w", "prim", "qf"):
            assert np.array_equal(ga[k], gb[k]), (s, k)
    a.close()
    b.close()


@pytest.mark.gpu
def test_graph_replay_follows_a_flux_waves_change():
    """Changing the flux kernel's workgroup size after graph steps drops the captured graphs:
    the replayed steps equal launches one by one with the same option change, bit for bit
    (the workgroup size fixes the order of the per-cell sums, so a stale graph would show)."""
    case = CASES["riemann16_amr_vsamr"]()
    say(f"[graph waves] cells={case.n_cells} P={case.phase_cells}")
    a, _ = pair(case, options={"flux_waves": 8})
    b, _ = pair(case, options={"flux_waves": 8, "graphs": 1})
    for s in range(6):
        if s == 3:
            a.set_option("flux_waves", 4)
            b.set_option("flux_waves", 4)
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
        ga, gb = a.get_state(), b.get_state()
        for k in ("df", "w", "prim", "qf"):
            assert np.array_equal(ga[k], gb[k]), (s, k)
    a.close()
    b.close()


@pytest.mark.gpu
def test_autotune_on_device_picks_a_timed_setting():
    """The autotuner times real steps on the device and returns one of the settings it
    timed, the fastest of them."""
    from kitamr_hip import autotune
    case = CASES["riemann8_amr1_nv16"]()
    say(f"[autotune] cells={case.n_cells} P={case.phase_cells}")
    best, ms, trials = autotune.autotune(case, steps=2, warmup=1, log=say)
    assert ms > 0 and best in [t[0] for t in trials]
    assert ms == min(t[1] for t in trials)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_walled_box_conserves_mass_on_device(precision):
    """Maxwellian walls on every side: the device's wall density keeps the total mass
    (FP64 and the FP32 form, whose w update stays FP64)."""
    case = cases.riemann_2d(n=8, nv=16, bc="Maxwellian")
    say(f"[parity] walled box mass, precision {precision}")
    dev, _ = pair(case, precision=precision)
    n = case.n_cells
    V = np.prod(np.asarray(case.topo["cell_ds"])[:n].reshape(n, 2), axis=1)
    m0 = (np.asarray(case.w).reshape(n, 4)[:, 0] * V).sum()
    for s in range(6):
        dev.step(case.dt_xi, s)
    m = (dev.get_state()["w"].reshape(n, 4)[:, 0] * V).sum()
    assert abs(m - m0) <= 1e-13 * m0, (m - m0) / m0
    dev.close()
