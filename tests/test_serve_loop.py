"""Host logic of the request loop (mctdhb_amd.serve.Server) with stand-in
instances, events and streams: admission only after arrival and within the
budget and the cap, every step enqueued in order (imaginary then real),
retirement only once the last step's event reports done, every step of a
request on the one stream it was given at admission, and a per-request
status that keeps failed solves out of the latency figures."""
import time

from mctdhb_amd import batch, serve


class FakeEvent:
    def __init__(self):
        self.recorded = False
        self.polls = 0

    def record(self, stream):
        self.recorded = True

    def query(self):
        # done on the second poll after recording (work takes a moment)
        if self.recorded:
            self.polls += 1
        return self.recorded and self.polls >= 2


class FakeInstance:
    log = []
    fail_rid = None

    def __init__(self, req):
        self.req = req
        self.steps = []
        FakeInstance.log.append(("admit", req.rid, time.perf_counter()))

    def step(self, stream=None):
        self.steps.append(("imag", stream))

    def step_real(self, stream=None):
        self.steps.append(("real", stream))

    def energy_per_particle(self):
        if self.req.rid == FakeInstance.fail_rid:
            return float("nan")
        return float(len(self.steps))

    def close(self):
        FakeInstance.log.append(("close", self.req.rid, time.perf_counter()))


def _requests():
    out = []
    for i, (t, kind, fp_dim) in enumerate([(0.0, "ground", 1000), (0.0, "quench", 1000), (0.01, "ground", 1000),
                                           (0.02, "quench", 50000), (0.03, "ground", 1000)]):
        d = batch.Draw(10, 3, 256, 1.0, 2.0, 0.5, i)
        q = serve.Request(i, t, kind, d, 5, 7 if kind == "quench" else 0, 1.5)
        out.append(q)
    return out


def test_loop_admits_within_budget_and_cap_and_runs_every_step():
    FakeInstance.log.clear()
    reqs = _requests()
    fps = [q.footprint for q in reqs]
    budget = sorted(fps)[-1] + sorted(fps)[-2]  # at most two at once
    instances = {}

    def make(req):
        inst = FakeInstance(req)
        instances[req.rid] = inst
        return inst

    srv = serve.Server(budget, nstreams=2, chunk=3, max_active=2, make_instance=make, make_event=FakeEvent,
                       make_stream=object, wait_all=lambda: None)
    recs = srv.run(reqs)
    assert sorted(r["rid"] for r in recs) == [0, 1, 2, 3, 4]
    assert srv.peak_active <= 2 and srv.peak_used <= budget
    for q in reqs:
        kinds = [k for k, _ in instances[q.rid].steps]
        assert kinds == ["imag"] * q.imag_steps + ["real"] * q.real_steps
        # one stream per request, whatever retirements do to the active list
        assert len({id(s) for _, s in instances[q.rid].steps}) == 1
    # requests were spread over both streams
    assert len({id(instances[q.rid].steps[0][1]) for q in reqs}) == 2
    for r in recs:
        assert r["status"] == "ok" and r["energy_per_particle"] == r["steps"]
        assert r["admitted"] >= r["arrival"] - 1e-9 and r["done"] >= r["admitted"]
        assert r["time_to_solution"] == r["done"] - r["arrival"]
    # every admission happened after its arrival and each instance was closed once
    t_admit = {rid: t for ev, rid, t in FakeInstance.log if ev == "admit"}
    closes = [rid for ev, rid, _ in FakeInstance.log if ev == "close"]
    assert sorted(closes) == [0, 1, 2, 3, 4] and len(t_admit) == 5


def test_failed_solves_are_marked_and_left_out_of_latency():
    FakeInstance.log.clear()
    FakeInstance.fail_rid = 3
    try:
        reqs = _requests()
        srv = serve.Server(10 * max(q.footprint for q in reqs), nstreams=3, chunk=2, make_instance=FakeInstance,
                           make_event=FakeEvent, make_stream=object, wait_all=lambda: None)
        recs = srv.run(reqs)
    finally:
        FakeInstance.fail_rid = None
    status = {r["rid"]: (r["status"], r["reason"]) for r in recs}
    assert status[3] == ("failed", "energy not finite")
    assert all(status[i] == ("ok", "") for i in (0, 1, 2, 4))
    lat = serve.latency_summary(recs)
    assert lat["count"] == 4
    ok = [r["time_to_solution"] for r in recs if r["rid"] != 3]
    assert lat["max_s"] == max(ok)


def test_loop_rejects_a_request_above_the_whole_budget():
    reqs = _requests()
    srv = serve.Server(min(q.footprint for q in reqs) - 1, make_instance=FakeInstance, make_event=FakeEvent,
                       make_stream=object, wait_all=lambda: None)
    try:
        srv.run(reqs)
    except ValueError as e:
        assert "above the whole budget" in str(e)
    else:
        raise AssertionError("expected a ValueError")


def test_retired_requests_are_released_in_bulk():
    FakeInstance.log.clear()
    reqs = _requests()
    big = 10 * max(q.footprint for q in reqs)
    srv = serve.Server(big, nstreams=2, chunk=4, make_instance=FakeInstance, make_event=FakeEvent,
                       make_stream=object, wait_all=lambda: None, reclaim_bytes=big)
    recs = srv.run(reqs)
    assert sorted(r["rid"] for r in recs) == [0, 1, 2, 3, 4]
    closes = [rid for ev, rid, _ in FakeInstance.log if ev == "close"]
    assert sorted(closes) == [0, 1, 2, 3, 4]
    assert srv.releases == 1  # nothing reached the threshold: one release at the end
    # every close after the last completion
    t_last_done = max(t for ev, rid, t in FakeInstance.log if ev == "admit")
    assert all(t >= t_last_done for ev, rid, t in FakeInstance.log if ev == "close")


def test_energy_bound_covers_the_grid_and_flags_divergence():
    """The divergence bound of Instance.health: above V(L) = 200 and the
    grid's largest kinetic energy, growing with |g| (N - 1), and using the
    larger of the relaxation and quench interactions."""
    import math

    d = batch.Draw(20, 3, 256, 1.0, 2.0, 0.5, 0)
    inst = batch.Instance.__new__(batch.Instance)
    inst.draw, inst.quench_g = d, None
    dx = 2 * d.half_width / d.ngrid
    b0 = inst.energy_bound()
    assert b0 >= 200.0 * (1 - 1e-12) + 0.5 * (math.pi / dx) ** 2
    inst.quench_g = -3.0  # an attractive quench, larger in magnitude
    assert inst.energy_bound() > b0
    assert 1.4e30 > inst.energy_bound()  # a diverged quench's E/N is far above it


def test_health_flags_a_rising_imaginary_time_energy():
    """An imaginary-time solve whose final energy is above its initial one
    fails (the split step was unstable); quenches are not judged by it."""
    import torch

    d = batch.Draw(2, 2, 256, 1.0, 2.0, 0.5, 0)
    inst = batch.Instance.__new__(batch.Instance)
    inst.draw, inst.quench_g = d, None
    e = 3.0
    inst.t = {"r1": torch.tensor([2.0, 0, 0, 0], dtype=torch.complex128),
              "hob": torch.tensor([e, 0, 0, e], dtype=torch.complex128),
              "r2": torch.zeros(16, dtype=torch.complex128), "hint": torch.zeros(16, dtype=torch.complex128),
              "coef": torch.tensor([1.0, 0, 0], dtype=torch.complex128)}
    assert inst.energy_per_particle() == e  # tr(rho1 h) / N
    inst.energy0 = e + 1.0
    assert inst.health() == ""
    inst.energy0 = e - 1.0
    assert inst.health() == "imaginary-time energy rose (unstable step)"
    inst.energy0 = e * (1 - 1e-3)  # a rise well below RELAX_RISE: not unstable
    assert inst.health() == ""
    inst.quench_g = 1.0
    assert inst.health() == ""  # a quench without snapshots: the final state only

    def snap(energy):  # rho1, h, rho2, g as serve's snapshots hold them, E/N = energy
        t = inst.t
        return torch.cat([t["r1"], torch.tensor([energy, 0, 0, energy], dtype=torch.complex128), t["r2"],
                          t["hint"]])

    inst.energy0 = e + 1.0
    # relaxed below E0 through the chunks; final within 10 % of the quench
    inst.snaps = [("imag", snap(e + 0.8)), ("relaxed", snap(e + 0.5)), ("quenched", snap(e * 1.05))]
    assert inst.health() == ""
    inst.snaps = [("imag", snap(e - 0.5)), ("relaxed", snap(e)), ("quenched", snap(e))]  # rose between chunks
    assert inst.health() == "imaginary-time energy rose (unstable step)"
    inst.snaps = [("relaxed", snap(e)), ("quenched", snap(2.0 * e))]  # final e against 2e after the quench
    assert inst.health() == "real-time energy not conserved (unstable step)"
    inst.quench_g = None  # a relaxation: its chunks, then the final state
    inst.snaps = [("imag", snap(e + 0.5))]
    assert inst.health() == ""
    inst.snaps = [("imag", snap(e - 0.5))]  # the final e is above the last chunk's
    assert inst.health() == "imaginary-time energy rose (unstable step)"
