"""The request loop (mctdhb_amd.serve): requests arrive over time, wait for
budget, run together, and each result is the same as alone (bit for bit)."""
import numpy as np
import pytest

from mctdhb_amd import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def test_served_results_equal_alone(amd):
    # seed 23: two quenches and four ground states among the first six
    reqs = serve.draw_requests(rate=300.0, window=0.03, seed=23, time_scale=0.012, dim_range=(1e2, 3e3))[:6]
    assert sum(q.kind == "quench" for q in reqs) == 2
    print("serve requests", [(q.kind, q.draw.npar, q.draw.norb, q.draw.dim, q.imag_steps, q.real_steps)
                             for q in reqs], flush=True)
    # room for the two largest only: later arrivals wait for a retirement
    fps = sorted(q.footprint for q in reqs)
    budget = fps[-1] + fps[-2]
    srv = serve.Server(budget_bytes=budget, nstreams=3, chunk=3)
    recs = srv.run(reqs, keep_state=True)
    assert sorted(r["rid"] for r in recs) == sorted(q.rid for q in reqs)
    assert srv.peak_used <= budget and srv.peak_active < len(reqs)
    by_id = {q.rid: q for q in reqs}
    for r in recs:
        alone = serve.run_alone(by_id[r["rid"]])
        for k in ("orb", "coef", "r1", "rinv"):
            assert np.array_equal(r["state"][k], alone["state"][k]), (r["rid"], k)
        assert r["time_to_solution"] >= 0 and r["done"] >= r["admitted"] >= r["arrival"] - 1e-9
        assert np.isfinite(r["energy_per_particle"])
        assert r["status"] == "ok" and r["reason"] == ""


def test_mid_size_requests_equal_alone_with_status(amd):
    """Requests of NC 1e4 - 1e5 (the mixed form) through the loop with a
    budget for two at a time: every record carries its status, and each
    result is bit-identical to the request run alone (the H.C tiling is fixed
    per (N, M) for the process)."""
    reqs = serve.draw_requests(rate=400.0, window=0.05, seed=31, time_scale=0.01, dim_range=(1e4, 1e5))[:5]
    assert len(reqs) == 5
    print("mid-size requests", [(q.kind, q.draw.npar, q.draw.norb, q.draw.dim, q.imag_steps, q.real_steps)
                                for q in reqs], flush=True)
    fps = sorted(q.footprint for q in reqs)
    srv = serve.Server(budget_bytes=fps[-1] + fps[-2], nstreams=2, chunk=2)
    recs = srv.run(reqs, keep_state=True)
    assert sorted(r["rid"] for r in recs) == sorted(q.rid for q in reqs)
    by_id = {q.rid: q for q in reqs}
    for r in recs:
        assert r["status"] in ("ok", "failed")
        alone = serve.run_alone(by_id[r["rid"]])
        for k in ("orb", "coef", "r1", "rinv"):  # a failed quench gives the same NaNs alone
            assert np.array_equal(r["state"][k], alone["state"][k], equal_nan=True), (r["rid"], k)
        ok = bool(np.isfinite(alone["energy_per_particle"]))
        assert (r["status"] == "ok") == ok or r["reason"] != "energy not finite"
