"""Autotuner search and cache (no GPU: a synthetic timer stands in for the device)."""
import json

from kitamr_hip import autotune, cases


def _timer(costs, calls):
    def t(opts):
        calls.append(dict(opts))
        return sum(costs[k][v] for k, v in opts.items())
    return t


def test_coordinate_descent_finds_the_separable_minimum():
    case = cases.riemann_2d(n=4, nv=16)
    costs = {"slope_vpt": {3: 3.0, 5: 1.0, 4: 2.0, 6: 4.0}, "flux_waves": {0: 0.8, 2: 0.9, 4: 0.5, 8: 1.0},
             "vel_tiles": {0: 2.0, 1: 1.0}, "branches": {1: 0.5, 2: 0.25, 4: 0.0}}
    calls = []
    best, ms, trials = autotune.autotune(case, timer=_timer(costs, calls))
    assert best == {"slope_vpt": 5, "flux_waves": 4, "vel_tiles": 1, "branches": 4}
    assert ms == 1.0 + 0.5 + 1.0
    assert len(trials) == len(calls) == 1 + 3 + 3 + 1 + 2     # start + the other values
    assert all("overlap_maps" not in c for c in calls)    # shared grid: no maps option


def test_adapted_grids_tune_maps_not_tiles():
    case = cases.riemann_2d(n=4, nv=8, vs_levels=1)
    space = autotune.relevant_space(case)
    assert "overlap_maps" in space and "vel_tiles" not in space


def test_cache_round_trip(tmp_path):
    case = cases.riemann_2d(n=4, nv=16)
    costs = {"slope_vpt": {3: 1.0, 4: 2.0, 5: 3.0, 6: 4.0}, "flux_waves": {0: 1.5, 2: 1.2, 4: 1.0, 8: 2.0},
             "vel_tiles": {0: 1.0, 1: 2.0}, "branches": {1: 1.0, 2: 0.5, 4: 0.75}}
    calls = []
    path = tmp_path / "tune.json"
    opts = autotune.tuned_options(case, cache_path=path, timer=_timer(costs, calls))
    assert opts == {"slope_vpt": 3, "flux_waves": 4, "vel_tiles": 0, "branches": 2}
    n = len(calls)
    again = autotune.tuned_options(case, cache_path=path, timer=_timer(costs, calls))
    assert again == opts and len(calls) == n               # served from the cache
    data = json.loads(path.read_text())
    assert autotune.shape_key(case) in data
