"""Serving loop: flow cases arrive over time, are admitted by HBM budget and slot count,
stay resident on the device and are stepped concurrently (continuous batching).

Each admitted case owns a device context with its own HIP stream (kah_create); a round of
the loop enqueues one step of every active case without waiting, so the kernels of
different cases overlap on the GPU. The host only blocks on one case's stream at that
case's residual-check steps (every ``ST_CHECK_INTERVAL`` steps, as the reference's solve
loop checks, src/Solver/Solver.jl:71-80). A case leaves when its residual is below the
solver's TOLERANCE for more than REDUNDANT_STEPS_NUM steps (check_for_convergence,
src/Solver/Finalize.jl:79-82) or at its step limit; its slot and memory are then free for
the next queued case (FIFO).

No counterpart in the reference (it runs one case per MPI job); this is the request loop
of the serving pattern, built on the same per-case API (``solver.KA``, ``solver.step``).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np


@dataclass
class Request:
    rid: int
    case: object                  # a prepared case (kitamr_hip.cases / solver.build_case)
    arrival: float                # seconds after the start of the run
    max_steps: int = 2000
    tag: str = ""
    nbytes: int = 0               # device bytes reserved for admission (an upper bound)
    charged: int = 0              # device bytes the created context holds (kah_device_bytes)
    admitted: Optional[float] = None
    finished: Optional[float] = None
    steps: int = 0
    converged: bool = False
    result: Optional[dict] = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def device_bytes_estimate(case):
    """Device bytes of one case before it is created: the f ping-pong buffers and the slopes
    (the bulk), velocity grids, per-cell state and descriptors (upper bounds)."""
    s = case.config.solver
    D, NDF = s.DIM, s.NDF
    fs = 4 if getattr(case.config, "precision", "fp64") == "fp32f" else 8
    P = int(np.asarray(case.topo["vs_offset"])[-1])
    nc = int(case.topo["n_cells"])
    npairs = int(case.topo["n_pairs"])
    state = P * NDF * fs * (2 + D)
    grids = (len(case.topo["vs_weight"]) * (D * 8 + 8 + 1 + 4))
    per_cell = nc * (8 * (4 * (D + 2) + 2 * (D + 2) + D + 2 * D * 3) + 2 * D * 96 + 8 * 32 * 160)
    halves = npairs * 2 * (D + 2) * 8 * max(1, P // max(nc, 1) // 256 + 1)
    scratch = P * (NDF + 1) * 8 if P // max(nc, 1) * (NDF + 1) * 8 > 64 * 1024 else 0
    # immersed boundary: the SolidNeighbor rows (f in both buffers, grid copies, slopes)
    ib = 0
    if len(case.topo.get("ib_face_pair", [])):
        from .abi import ib_sizes
        rows = int(ib_sizes(case.topo).sum())
        ib = rows * (NDF * (2 * fs + 8) + D * 8 + 8 + 1 + 4) + len(case.topo["ib_face_pair"]) * 256
    return int(state + grids + per_cell + halves + scratch + ib + (1 << 20))


def _default_factory(case, options=None):
    from .solver import KA
    return KA(case, options=options)


class Server:
    """Continuous-batching request loop over device-resident cases.

    budget_bytes: device memory the resident cases may hold together.
    max_slots:    cases stepped concurrently (one stream each).
    factory:      case -> KA-like handle with ``.ctx`` (``step``, ``sync``, ``get_state``,
                  ``close``) and ``.status``; the default creates a device context with
                  ``options`` (launch options; by default each case's step is replayed
                  from captured graphs).
    """

    def __init__(self, budget_bytes, max_slots=8, factory: Callable = None,
                 clock: Callable[[], float] = time.perf_counter, log=None, options=None):
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        opts = {"graphs": 1} if options is None else options
        self.factory = factory or (lambda case: _default_factory(case, opts))
        self.clock = clock
        self.log = log
        self.used = 0
        self.peak_used = 0
        self.peak_active = 0
        self.admit_order = []

    def _say(self, msg):
        if self.log:
            self.log(msg)

    def run(self, requests, keep_df=True):
        from . import solver as S
        for r in requests:
            if not r.nbytes:
                r.nbytes = device_bytes_estimate(r.case)
            if r.nbytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.nbytes} B, over the budget "
                                 f"{self.budget} B")
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue = collections.deque()
        active = []
        t0 = self.clock()
        updates = 0
        while pending or queue or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # admission: FIFO, while a slot and the memory for the head request are free;
            # the request is admitted on its upper-bound estimate, then charged what its
            # context actually holds on the device
            while (queue and len(active) < self.max_slots
                   and self.used + queue[0].nbytes <= self.budget):
                r = queue.popleft()
                r.ka = self.factory(r.case)
                r.admitted = self.clock() - t0
                held = getattr(r.ka.ctx, "device_bytes", None)
                r.charged = int(held()) if held else r.nbytes
                self.used += r.charged
                self.peak_used = max(self.peak_used, self.used)
                active.append(r)
                self.admit_order.append(r.rid)
                self._say(f"[serve] t={r.admitted:.3f}s admit {r.rid} {r.tag} "
                          f"({r.charged / 1e6:.1f} MB, {len(active)} active)")
            self.peak_active = max(self.peak_active, len(active))
            if not active:
                if pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            # one step of every active case, enqueued on its own stream
            for r in active:
                S.limit_dt(r.ka)
                S.step(r.ka)
                r.steps += 1
            done = []
            for r in active:
                conv = S.check_for_convergence(r.ka)
                if conv or r.steps >= r.max_steps or S.reached_max_time(r.ka):
                    r.converged = bool(conv)
                    done.append(r)
            for r in done:
                r.ka.ctx.sync()
                st = r.ka.ctx.get_state(df=keep_df)
                r.result = st
                r.finished = self.clock() - t0
                updates += int(r.case.phase_cells) * r.steps
                r.ka.close()
                r.ka = None
                self.used -= r.charged
                active.remove(r)
                self._say(f"[serve] t={r.finished:.3f}s done {r.rid} {r.tag} steps={r.steps} "
                          f"latency={r.latency:.3f}s")
        wall = self.clock() - t0
        lat = np.array([r.latency for r in requests])
        return {
            "requests": len(requests),
            "wall_s": wall,
            "updates": updates,
            "updates_per_s": updates / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "peak_active": self.peak_active,
            "peak_bytes": self.peak_used,
            "budget_bytes": self.budget,
        }


def poisson_arrivals(n, rate, seed):
    """Arrival times (s) of n requests of a Poisson process of the given rate (1/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def c5_specs(n_req, seed):
    """Requests of the serving workload, drawn (seeded) from its stated distributions: case
    type uniform over {Riemann, cylinder}; Kn = 10^U(-3, 0); Ma ~ U(0.5, 5); physical cells
    log-uniform in [4k, 200k] (an n x n mesh, n = round(sqrt(cells)) rounded to even); velocity
    grids m x m with m uniform over {16, 24, ..., 64} on [-V, V]^2, V = max(5, Ma sqrt(5/6) + 5)
    (m even, so v = 0 lies on a cell corner); FP64 or (one in four) the FP32 form."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        kind = "riemann" if rng.random() < 0.5 else "cylinder"
        Kn = float(10.0 ** rng.uniform(-3.0, 0.0))
        Ma = float(rng.uniform(0.5, 5.0))
        cells = float(np.exp(rng.uniform(np.log(4e3), np.log(2e5))))
        n = max(64, 2 * int(round(np.sqrt(cells) / 2.0)))
        m = int(rng.choice(np.arange(16, 65, 8)))
        vmax = max(5.0, Ma * np.sqrt(5.0 / 6.0) + 5.0)
        prec = "fp32f" if rng.random() < 0.25 else "fp64"
        out.append(dict(kind=kind, Kn=Kn, Ma=Ma, n=n, nv=m, vmax=vmax, precision=prec))
    return out


def c5_case(spec):
    """The case of one serving request (c5_specs): the four-quadrant Riemann problem on
    [0,1]^2 with outflow boundaries, or Circle(Maxwellian, r = 1) in [-4, 4]^2 with a Mach-Ma
    inflow, at the reference's default CFL 0.4."""
    from . import cases
    if spec["kind"] == "riemann":
        return cases.riemann_2d(n=spec["n"], nv=spec["nv"], vmax=spec["vmax"], Kn=spec["Kn"],
                                CFL=0.4, precision=spec["precision"])
    return cases.cylinder_2d(n=spec["n"], nv=spec["nv"], vmax=spec["vmax"], Ma=spec["Ma"],
                             Kn=spec["Kn"], CFL=0.4, precision=spec["precision"])


def c5_tag(spec):
    return (f"{spec['kind']} {spec['n']}^2 cells {spec['nv']}^2 vel Kn={spec['Kn']:.2g} "
            f"Ma={spec['Ma']:.2f} {spec['precision']}")
