"""The benchmark's contract without a GPU: the algorithmic bytes per phase-space update that
its roofline figures use (per phase of the reference's dataflow: slope R df + W sdf; flux R df,
R sdf, W flux; iterate R df, R flux, W df; plus the velocity-grid geometry read by each phase
on per-cell adapted grids), and the refusal of more than one GPU."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


@pytest.mark.parametrize("s_f,adaptive,dim,ndf,want", [
    (4, True, 2, 2, (25, 105, 130)),     # the default line: 2D2F FP32 form, adapted grids
    (8, True, 2, 2, (49, 161, 210)),     # its FP64 form
    (8, False, 2, 2, (48, 112, 160)),    # shared grid, FP64
    (4, False, 2, 2, (24, 56, 80)),      # shared grid, FP32 form
    (8, False, 3, 1, (32, 64, 96)),      # 3D1F shared grid, FP64
    (8, True, 3, 1, (33, 129, 162)),     # 3D1F adapted grids, FP64
    (4, True, 3, 1, (17, 97, 114)),      # 3D1F adapted grids, FP32 form
])
def test_algorithmic_bytes_per_update(s_f, adaptive, dim, ndf, want):
    slope, flux_iter, step = bench.alg_bytes(s_f, adaptive, dim, ndf)
    assert (slope, flux_iter, step) == want
    # the step is the sum of its phases; the slope phase reads f (ndf values) and writes
    # dim slopes of each
    assert step == slope + flux_iter
    assert slope - (1 if adaptive else 0) == ndf * s_f * (1 + dim)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
