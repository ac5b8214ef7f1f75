"""Restart checkpoints of the device state: the per-rank phase-space arrays of
``save_for_restart`` / ``restart`` (src/IO/Restart.jl:73-431).

The arrays (``vs_nums``, ``bound_encs``, ``ws``, ``vs_levels``, ``vs_midpoints``, ``vs_df``)
are gathered from and scattered to HBM by the library (``kah_save_restart`` /
``kah_load_restart``); this module is the host container around them:

* ``restart_<rank>.npz`` holds the six arrays under the reference's names and logical shapes
  (the reference writes them into ``restart_<rank>.jld2``; no HDF5/JLD2 writer is available to
  this package, so the container is numpy's, the arrays are the same);
* ``manifest.json`` holds format_version, DIM, NDF, mpi_size, global_quad_num, gfq,
  local_quad_nums and vs_totals (src/IO/Restart.jl:176-185);
* ``status.json`` holds the time-loop scalars (src/IO/Restart.jl:165-174).

The p4est forest files and the configuration belong to the host's forest and are not written
here. Loading follows ``_restart_build_ps_list`` (src/IO/Restart.jl:258-323): the rank files
that overlap this rank's global cell range are concatenated and sliced, so a checkpoint
written by one partition loads on another.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

RESTART_FORMAT_VERSION = 1   # src/IO/Restart.jl:17
ARRAYS = ("vs_nums", "bound_encs", "ws", "vs_levels", "vs_midpoints", "vs_df")
STATUS_KEYS = ("step", "sim_time", "Dt", "Dt_xi", "gradmax", "ps_adapt_step", "vs_adapt_step",
               "partition_step")


def _dir(path) -> Path:
    p = Path(path)
    p.mkdir(parents=True, exist_ok=True)
    return p


def write_rank_file(dir_path, rank, arrays):
    """restart_<rank>.npz with the six per-rank arrays."""
    missing = [k for k in ARRAYS if k not in arrays]
    if missing:
        raise ValueError(f"restart arrays missing: {missing}")
    np.savez_compressed(_dir(dir_path) / f"restart_{rank}.npz",
                        **{k: np.asarray(arrays[k]) for k in ARRAYS})


def write_manifest(dir_path, dim, ndf, local_quad_nums, vs_totals):
    lqn = [int(x) for x in local_quad_nums]
    gfq = [0]
    for n in lqn:
        gfq.append(gfq[-1] + n)
    m = dict(format_version=RESTART_FORMAT_VERSION, DIM=int(dim), NDF=int(ndf),
             mpi_size=len(lqn), global_quad_num=gfq[-1], gfq=gfq, local_quad_nums=lqn,
             vs_totals=[int(x) for x in vs_totals])
    (_dir(dir_path) / "manifest.json").write_text(json.dumps(m, indent=1))
    return m


def write_status(dir_path, status=None):
    st = {k: 0 for k in STATUS_KEYS}
    st.update(status or {})
    (_dir(dir_path) / "status.json").write_text(json.dumps(st, indent=1))


def save_for_restart(dev, dir_path="restart", status=None):
    """One-rank checkpoint of a context's state (src/IO/Restart.jl:73-209); returns the path."""
    a = dev.save_restart()
    write_rank_file(dir_path, 0, a)
    write_manifest(dir_path, dev.dim, dev.ndf, [len(a["vs_nums"])], [int(a["vs_nums"].sum())])
    write_status(dir_path, status)
    return str(dir_path)


def integrity_message(dir_path):
    """'' when the manifest, status and per-rank files agree, else the list of problems
    (_restart_integrity_message, src/IO/Restart.jl:233-256)."""
    d = Path(dir_path)
    problems = []
    mf = d / "manifest.json"
    if not mf.is_file():
        return f"Restart integrity check failed:\n  - missing manifest: {mf}"
    m = json.loads(mf.read_text())
    if m.get("format_version") != RESTART_FORMAT_VERSION:
        problems.append(f"manifest format_version {m.get('format_version')} != expected "
                        f"{RESTART_FORMAT_VERSION}")
    if not (d / "status.json").is_file():
        problems.append(f"missing metadata file: {d / 'status.json'}")
    for r in range(int(m["mpi_size"])):
        f = d / f"restart_{r}.npz"
        if not f.is_file():
            problems.append(f"missing per-rank data file: {f}")
            continue
        try:
            with np.load(f) as z:
                vsn = z["vs_nums"]
        except Exception as e:   # unreadable file: reported like the reference's catch
            problems.append(f"{f}: failed to read ({e})")
            continue
        if len(vsn) != m["local_quad_nums"][r]:
            problems.append(f"{f}: cell count {len(vsn)} != manifest {m['local_quad_nums'][r]}")
        if int(vsn.sum()) != m["vs_totals"][r]:
            problems.append(f"{f}: velocity-cell total {int(vsn.sum())} != manifest "
                            f"{m['vs_totals'][r]}")
    return "" if not problems else "Restart integrity check failed:\n  - " + "\n  - ".join(problems)


def load_cells(dir_path, g0, g1):
    """Restart arrays of the global cells [g0, g1): the overlapping rank files concatenated and
    sliced (src/IO/Restart.jl:266-300)."""
    d = Path(dir_path)
    m = json.loads((d / "manifest.json").read_text())
    gfq = np.asarray(m["gfq"], dtype=np.int64)
    dim, ndf = int(m["DIM"]), int(m["NDF"])
    out = {k: [] for k in ARRAYS}
    if g1 <= g0:
        return {"vs_nums": np.zeros(0, np.int64), "bound_encs": np.zeros(0, np.int64),
                "ws": np.zeros((0, dim + 2)), "vs_levels": np.zeros(0, np.int8),
                "vs_midpoints": np.zeros((0, dim)), "vs_df": np.zeros((0, ndf))}
    if g0 < 0 or g1 > gfq[-1]:
        raise ValueError(f"cells [{g0}, {g1}) outside the checkpoint's {gfq[-1]} cells")
    first = int(np.searchsorted(gfq, g0, side="right")) - 1
    last = int(np.searchsorted(gfq, g1 - 1, side="right")) - 1
    for r in range(first, last + 1):
        with np.load(d / f"restart_{r}.npz") as z:
            for k in ARRAYS:
                out[k].append(z[k])
    cat = {k: np.concatenate(v) for k, v in out.items()}
    sa = g0 - int(gfq[first])
    cnt = g1 - g0
    cum = np.concatenate([[0], np.cumsum(cat["vs_nums"])])
    v0, v1 = int(cum[sa]), int(cum[sa + cnt])
    return {"vs_nums": cat["vs_nums"][sa:sa + cnt], "bound_encs": cat["bound_encs"][sa:sa + cnt],
            "ws": cat["ws"][sa:sa + cnt], "vs_levels": cat["vs_levels"][v0:v1],
            "vs_midpoints": cat["vs_midpoints"][v0:v1], "vs_df": cat["vs_df"][v0:v1]}


def read_status(dir_path):
    return json.loads((Path(dir_path) / "status.json").read_text())


def restart(dev, dir_path, g0=0, g1=None, check_integrity=True):
    """Loads the checkpoint's cells [g0, g1) (default: all) into a context whose topology is
    set, and returns the status scalars (src/IO/Restart.jl:350-431)."""
    if check_integrity:
        msg = integrity_message(dir_path)
        if msg:
            raise RuntimeError(msg)
    m = json.loads((Path(dir_path) / "manifest.json").read_text())
    if (int(m["DIM"]), int(m["NDF"])) != (dev.dim, dev.ndf):
        raise RuntimeError(f"checkpoint is {m['DIM']}D/NDF={m['NDF']}; the context is "
                           f"{dev.dim}D/NDF={dev.ndf}")
    a = load_cells(dir_path, g0, m["global_quad_num"] if g1 is None else g1)
    dev.load_restart(a)
    return read_status(dir_path)
