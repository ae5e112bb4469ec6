"""Request draws of mctdhb_amd.serve (host side only)."""
import math

from mctdhb_amd import serve


def test_npar_for_dim_is_smallest():
    for m in (2, 3, 5, 7):
        for target in (100, 5000, 123456, 5e6):
            n = serve.npar_for_dim(target, m, max_npar=4000)
            assert n == 4000 or math.comb(n + m - 1, m - 1) >= target
            assert n == 1 or math.comb(n - 1 + m - 1, m - 1) < target


def test_request_mix():
    rq = serve.draw_requests(200.0, 2.0, seed=3)
    assert 300 < len(rq) < 500  # Poisson, mean 400
    assert all(0 <= a.arrival < b.arrival < 2.0 for a, b in zip(rq, rq[1:]))
    q = sum(r.kind == "quench" for r in rq) / len(rq)
    assert 0.2 < q < 0.4
    for r in rq:
        assert 2 <= r.draw.norb <= 7 and r.draw.ngrid in (512, 1024, 2048)
        assert r.draw.dim <= 6e6 and r.imag_steps in (200, 500)
        if r.kind == "quench":
            assert 200 <= r.real_steps <= 2000 and 0.5 <= r.g_factor <= 2.0
        assert r.footprint > 16 * r.draw.dim
