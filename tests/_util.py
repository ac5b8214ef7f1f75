"""Shared helpers of the parity tests: run the device path and the oracle side by side."""
import sys

import numpy as np

import oracle as O
from kitamr_hip import abi
from kitamr_hip.solver import kah_params


def rel_err(a, b):
    """max |a-b| / max |b| over an array (scale-aware: components near zero do not blow up)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def rel_err_rows(a, b, width):
    """Per-row (per cell) max |a-b| / max |b|, reduced by max: each cell measured on its own scale."""
    a = np.asarray(a).reshape(-1, width)
    b = np.asarray(b).reshape(-1, width)
    scale = np.maximum(np.max(np.abs(b), axis=1), 1e-300)
    return float(np.max(np.max(np.abs(a - b), axis=1) / scale))


def pair(case, options=None, **kw):
    params = kah_params(case.config)
    for k, v in kw.items():
        setattr(params, k, v)
    dev = abi.HipContext(params)
    for k, v in (options or {}).items():
        dev.set_option(k, v)
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    orc = O.Oracle(params, case.topo)
    orc.set_state(case.w, case.prim, case.df)
    return dev, orc


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err_cols(a, b, width):
    """Per component (column) max |a-b| / max |b|, reduced by max. For face-flux sums that
    cancel to ~0 inside uniform regions the column scale stands in for the sum of |terms|."""
    a = np.asarray(a).reshape(-1, width)
    b = np.asarray(b).reshape(-1, width)
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return float(np.max(np.max(np.abs(a - b), axis=0) / scale))
