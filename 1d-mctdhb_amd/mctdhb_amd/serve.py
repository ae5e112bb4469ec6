"""A request loop over many MCTDHB solves on one GPU (continuous batching).

Requests arrive over time.  A request is admitted as soon as it has arrived
and its device footprint fits in what is left of a fixed HBM budget; from
then on its time steps are replayed (captured graphs, mctdhb_amd.batch)
together with those of every other admitted request, a few steps per
request per pass, round-robin over a few streams.  A request that has
enqueued its last step is retired when the event recorded after that step
has completed: its result is read, its memory returned to the budget, and
its time to solution (completion minus arrival) recorded.  Requests share no
buffers, so each result is the same as when the request runs alone.

Two kinds, as in the served mix: a ground state (imaginary-time steps) and
an interaction quench (imaginary-time relaxation, then g scaled by a factor
and real-time steps to a final time).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import batch, inputs


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the loop
    kind: str               # "ground" or "quench"
    draw: batch.Draw
    imag_steps: int
    real_steps: int = 0
    g_factor: float = 1.0

    @property
    def footprint(self) -> int:
        return batch.footprint_bytes(self.draw.dim, self.draw.norb, self.draw.ngrid + 1)


def npar_for_dim(target: float, norb: int, max_npar: int = 10_000) -> int:
    """Smallest N with C(N + M - 1, M - 1) >= target (at most max_npar)."""
    lo, hi = 1, max_npar
    while lo < hi:
        mid = (lo + hi) // 2
        if math.comb(mid + norb - 1, norb - 1) >= target:
            hi = mid
        else:
            lo = mid + 1
    return lo


def draw_requests(rate: float, window: float, seed: int = 13, time_scale: float = 1.0, dt: float = 0.005,
                  dim_range=(1e2, 5e6)) -> list[Request]:
    """Poisson arrivals at `rate` per second during `window` seconds; 70 %
    ground states (500 imaginary-time steps), 30 % quenches (200
    imaginary-time steps, then g times U[0.5, 2] and real time to
    T ~ U[1, 10]); NC log-uniform in dim_range with M in {2..7} and N chosen
    to reach it (N <= 4000: the fused H.C kernel keeps rank and square-root
    tables of N + 5 entries in LDS); Ng in {512, 1024, 2048}; traps and g as in
    batch.draw_ground_states.  time_scale multiplies every step count (a
    shortened sample keeps the same mix)."""
    r = inputs.rng(seed, 5)
    out, t = [], 0.0
    while True:
        t += float(r.exponential(1.0 / rate))
        if t >= window:
            return out
        kind = "ground" if r.random() < 0.7 else "quench"
        target = math.exp(r.uniform(math.log(dim_range[0]), math.log(dim_range[1])))
        m = int(r.integers(2, 8))
        n = npar_for_dim(target, m, max_npar=4000)
        ng = int(r.choice([512, 1024, 2048]))
        a, b = float(r.uniform(0.5, 2.0)), float(r.uniform(1.0, 3.0))
        g = float(10 ** r.uniform(-2.0, 1.0))
        f, T = float(r.uniform(0.5, 2.0)), float(r.uniform(1.0, 10.0))
        d = batch.Draw(n, m, ng, a, b, g, seed * 1000003 + len(out))
        if kind == "ground":
            req = Request(len(out), t, kind, d, max(1, round(500 * time_scale)))
        else:
            req = Request(len(out), t, kind, d, max(1, round(200 * time_scale)),
                          max(1, round(T / dt * time_scale)), f)
        out.append(req)


def _gpu_instance(req: Request):
    quench_g = req.draw.g * req.g_factor if req.kind == "quench" else None
    return batch.Instance(req.draw, quench_g=quench_g)


def _gpu_event():
    import torch

    return torch.cuda.Event()


def _gpu_stream():
    import torch

    return torch.cuda.Stream()


def _gpu_wait_all():
    import torch

    torch.cuda.synchronize()


class _Active:
    """An admitted request.  All of its steps and its completion event go on
    the one stream it is given at admission, so they run in order without
    any cross-stream dependency."""

    def __init__(self, req: Request, admitted: float, make_instance, make_event, stream):
        self.req = req
        self.admitted = admitted
        self.inst = make_instance(req)
        self.stream = stream
        self.done_steps = 0
        self.total = req.imag_steps + req.real_steps
        self.event = make_event()
        self.enqueued_all = False
        self.next_snap = self.SNAP_EVERY

    # relaxation steps between two snapshots
    SNAP_EVERY = 32

    def advance(self, k: int):
        # snapshots for health(), copied on this request's stream without
        # waiting: every SNAP_EVERY relaxation steps, and the relaxed state
        # and the state one real-time step after a quench
        snap = getattr(self.inst, "snapshot", None)
        for _ in range(min(k, self.total - self.done_steps)):
            if self.done_steps < self.req.imag_steps:
                self.inst.step(self.stream)
                if snap is not None and self.done_steps + 1 == self.next_snap < self.req.imag_steps:
                    snap(self.stream, "imag")
                    self.next_snap += self.SNAP_EVERY
            else:
                first = self.done_steps == self.req.imag_steps
                if first and snap is not None:
                    snap(self.stream, "relaxed")
                self.inst.step_real(self.stream)
                if first and snap is not None:
                    snap(self.stream, "quenched")
            self.done_steps += 1
        if self.done_steps == self.total and not self.enqueued_all:
            self.event.record(self.stream)
            self.enqueued_all = True


def result_status(inst) -> tuple[str, str]:
    """("ok", "") when the result is usable: finite energy, finite rho1 whose
    trace is N, and coefficients of unit norm; else ("failed", reason)."""
    e = inst.energy_per_particle()
    if not math.isfinite(e):
        return "failed", "energy not finite"
    check = getattr(inst, "health", None)
    if check is not None:
        reason = check()
        if reason:
            return "failed", reason
    return "ok", ""


class Server:
    """budget_bytes: HBM for the admitted requests' state; chunk: steps per
    request per pass; max_active: at most this many admitted at once (None:
    only the budget limits), which trades sharing for earlier completions."""

    def __init__(self, budget_bytes: int, nstreams: int = 4, chunk: int = 8, max_active: int | None = None,
                 make_instance=_gpu_instance, make_event=_gpu_event, make_stream=_gpu_stream,
                 wait_all=_gpu_wait_all, reclaim_bytes: int | None = None):
        """make_instance / make_event / make_stream / wait_all: the GPU
        objects by default; a test of the loop's host logic passes stand-ins.
        reclaim_bytes: a retired request's memory returns to the budget at
        once, but its device buffers are released in bulk once this many
        bytes of retired requests are held (and at the end): releasing
        device memory waits for the whole device, so doing it per retirement
        would drain the GPU each time.  None releases at every retirement."""
        self.budget = budget_bytes
        self.reclaim_bytes = reclaim_bytes
        self.streams = [make_stream() for _ in range(max(1, nstreams))]
        self.chunk = chunk
        self.max_active = max_active
        self._make_instance, self._make_event, self._wait_all = make_instance, make_event, wait_all

    def run(self, requests: list[Request], keep_state: bool = False) -> list[dict]:
        """Serve `requests` (arrival times relative to the call); returns one
        record per request, in completion order."""
        pending = deque(sorted(requests, key=lambda q: q.arrival))
        for q in pending:
            if q.footprint > self.budget:
                raise ValueError(f"request {q.rid} needs {q.footprint} bytes, above the whole budget")
        active: list[_Active] = []
        used, peak_active, peak_used, out = 0, 0, 0, []
        retired, retired_bytes, self.releases = [], 0, 0

        def release_retired():
            nonlocal retired, retired_bytes
            if retired:
                for inst in retired:
                    inst.close()
                self.releases += 1
            retired, retired_bytes = [], 0

        admitted = 0  # round-robin stream assignment at admission
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            while (pending and pending[0].arrival <= now and used + pending[0].footprint <= self.budget
                   and (self.max_active is None or len(active) < self.max_active)):
                q = pending.popleft()
                stream = self.streams[admitted % len(self.streams)]
                admitted += 1
                active.append(_Active(q, time.perf_counter() - t0, self._make_instance, self._make_event, stream))
                used += q.footprint
                peak_used = max(peak_used, used)
            peak_active = max(peak_active, len(active))
            if not active:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            for a in active:
                if not a.enqueued_all:
                    a.advance(self.chunk)
            still = []
            for a in active:
                if a.enqueued_all and a.event.query():
                    done = time.perf_counter() - t0
                    status, reason = result_status(a.inst)
                    obs = a.inst.observables() if hasattr(a.inst, "observables") else None
                    rec = {"rid": a.req.rid, "kind": a.req.kind, "dim": a.req.draw.dim, "norb": a.req.draw.norb,
                           "npar": a.req.draw.npar, "steps": a.total, "arrival": a.req.arrival,
                           "admitted": a.admitted, "done": done, "time_to_solution": done - a.req.arrival,
                           "energy_per_particle": obs[0] if obs is not None else a.inst.energy_per_particle(),
                           "status": status,
                           "reason": reason}
                    rec["mixed"] = bool(getattr(a.inst, "prec", 0))
                    if obs is not None:
                        rec["observables"] = obs
                    if keep_state:
                        rec["state"] = {k: v.cpu().numpy().copy() for k, v in a.inst.t.items()}
                    out.append(rec)
                    used -= a.req.footprint
                    if self.reclaim_bytes is None:
                        a.inst.close()
                    else:
                        retired.append(a.inst)
                        retired_bytes += a.req.footprint
                        if retired_bytes >= self.reclaim_bytes:
                            release_retired()
                else:
                    still.append(a)
            if len(still) == len(active) and all(a.enqueued_all for a in active):
                self._wait_all()  # only completions are left to wait for
            active = still
        release_retired()
        self.peak_active, self.peak_used = peak_active, peak_used
        return out


def run_alone(req: Request, fp64: bool = False) -> dict:
    """The same request on its own (default stream), for comparisons; fp64:
    its FP64 twin, whatever the size."""
    import torch

    quench_g = req.draw.g * req.g_factor if req.kind == "quench" else None
    inst = batch.Instance(req.draw, quench_g=quench_g, fp64=fp64)
    for _ in range(req.imag_steps):
        inst.step()
    for _ in range(req.real_steps):
        inst.step_real()
    torch.cuda.synchronize()
    st = {k: v.cpu().numpy().copy() for k, v in inst.t.items()}
    e = inst.energy_per_particle()
    obs = inst.observables()
    inst.close()
    return {"state": st, "energy_per_particle": e, "observables": obs}


def attach_fp64_errors(records: list[dict], requests: list[Request]) -> None:
    """For every record of a mixed-precision solve, run the request's FP64
    twin alone and store the relative errors of its energy, natural
    occupations and rho1 (`errors_vs_fp64`); FP64 solves get zeros."""
    by_id = {q.rid: q for q in requests}
    for r in records:
        if not r.get("mixed"):
            r["errors_vs_fp64"] = {"energy_rel": 0.0, "natural_occupations_rel": 0.0,
                                   "natural_occupations_abs": 0.0, "rho1_rel_fro": 0.0}
            continue
        twin = run_alone(by_id[r["rid"]], fp64=True)
        r["errors_vs_fp64"] = batch.errors_vs_fp64(r["observables"], twin["observables"])


def latency_summary(records: list[dict]) -> dict:
    """Time to solution over the requests that solved (status "ok")."""
    tts = np.array([r["time_to_solution"] for r in records if r.get("status", "ok") == "ok"])
    if tts.size == 0:
        return {"p50_s": None, "p99_s": None, "max_s": None, "count": 0}
    return {"p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "max_s": float(tts.max()), "count": int(tts.size)}
