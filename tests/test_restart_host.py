"""Host container of restart checkpoints (kitamr_hip.restart): the per-rank files are sliced
and concatenated like _restart_build_ps_list (src/IO/Restart.jl:258-323), so a checkpoint
written by one partition loads on another; the integrity check reports what
_restart_integrity_message (:233-256) reports. Pinned against direct slicing of the arrays."""
import json

import numpy as np
import pytest

from kitamr_hip import restart as R


def _arrays(n, dim=2, ndf=2, seed=0):
    rng = np.random.default_rng(seed)
    vs_nums = rng.integers(3, 9, size=n).astype(np.int64)
    solid = rng.random(n) < 0.2
    vs_nums[solid] = 0
    benc = np.where(solid, -1, 0).astype(np.int64)
    nv = int(vs_nums.sum())
    ws = rng.standard_normal((n, dim + 2))
    ws[solid] = 0.0
    return dict(vs_nums=vs_nums, bound_encs=benc, ws=ws,
                vs_levels=rng.integers(0, 3, size=nv).astype(np.int8),
                vs_midpoints=rng.standard_normal((nv, dim)), vs_df=rng.standard_normal((nv, ndf)))


def _write_partition(d, a, cuts, dim=2, ndf=2):
    cum = np.concatenate([[0], np.cumsum(a["vs_nums"])])
    lqn, vst = [], []
    for r in range(len(cuts) - 1):
        c0, c1 = cuts[r], cuts[r + 1]
        v0, v1 = int(cum[c0]), int(cum[c1])
        R.write_rank_file(d, r, {"vs_nums": a["vs_nums"][c0:c1],
                                 "bound_encs": a["bound_encs"][c0:c1], "ws": a["ws"][c0:c1],
                                 "vs_levels": a["vs_levels"][v0:v1],
                                 "vs_midpoints": a["vs_midpoints"][v0:v1],
                                 "vs_df": a["vs_df"][v0:v1]})
        lqn.append(c1 - c0)
        vst.append(v1 - v0)
    R.write_manifest(d, dim, ndf, lqn, vst)
    R.write_status(d, {"step": 17, "sim_time": 0.125})


def _expect(a, g0, g1):
    cum = np.concatenate([[0], np.cumsum(a["vs_nums"])])
    v0, v1 = int(cum[g0]), int(cum[g1])
    return {"vs_nums": a["vs_nums"][g0:g1], "bound_encs": a["bound_encs"][g0:g1],
            "ws": a["ws"][g0:g1], "vs_levels": a["vs_levels"][v0:v1],
            "vs_midpoints": a["vs_midpoints"][v0:v1], "vs_df": a["vs_df"][v0:v1]}


@pytest.mark.parametrize("cuts", [[0, 23], [0, 10, 23], [0, 4, 4, 15, 23], [0, 0, 23, 23]])
def test_any_new_partition_reads_the_same_cells(tmp_path, cuts):
    a = _arrays(23)
    _write_partition(tmp_path, a, cuts)
    assert R.integrity_message(tmp_path) == ""
    for new in ([0, 23], [0, 11, 23], [0, 3, 9, 16, 23], [0, 1, 2, 22, 23]):
        for r in range(len(new) - 1):
            got = R.load_cells(tmp_path, new[r], new[r + 1])
            exp = _expect(a, new[r], new[r + 1])
            for k in R.ARRAYS:
                assert np.array_equal(got[k], exp[k]), (cuts, new, r, k)
    assert R.read_status(tmp_path)["step"] == 17


def test_integrity_reports_missing_and_inconsistent_files(tmp_path):
    a = _arrays(12)
    _write_partition(tmp_path, a, [0, 5, 12])
    (tmp_path / "restart_1.npz").unlink()
    m = json.loads((tmp_path / "manifest.json").read_text())
    m["vs_totals"][0] += 1
    m["format_version"] = 99
    (tmp_path / "manifest.json").write_text(json.dumps(m))
    msg = R.integrity_message(tmp_path)
    assert "missing per-rank data file" in msg and "restart_1.npz" in msg
    assert "velocity-cell total" in msg
    assert "format_version 99" in msg
    (tmp_path / "status.json").unlink()
    assert "missing metadata file" in R.integrity_message(tmp_path)


def test_out_of_range_cells_raise(tmp_path):
    a = _arrays(8)
    _write_partition(tmp_path, a, [0, 8])
    with pytest.raises(ValueError):
        R.load_cells(tmp_path, 2, 9)
    empty = R.load_cells(tmp_path, 3, 3)
    assert all(len(v) == 0 for v in empty.values())
