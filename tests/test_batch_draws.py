"""The uneven ground-state draws of mctdhb_amd.batch (host side only)."""
import math

from mctdhb_amd import batch


def test_draws_respect_ranges_and_cap():
    ds = batch.draw_ground_states(300, seed=2, max_dim=200_000)
    assert len(ds) == 300
    for d in ds:
        assert 4 <= d.npar <= 64 and 2 <= d.norb <= 6
        assert d.ngrid in (256, 512, 1024)
        assert 0.5 <= d.a <= 2.0 and 1.0 <= d.b <= 3.0 and 0.01 <= d.g <= 10.0
        assert d.dim == math.comb(d.npar + d.norb - 1, d.norb - 1) <= 200_000
    assert len({d.norb for d in ds}) == 5 and len({d.ngrid for d in ds}) == 3


def test_draws_are_reproducible():
    a = batch.draw_ground_states(20, seed=4)
    b = batch.draw_ground_states(20, seed=4)
    assert a == b
    assert a != batch.draw_ground_states(20, seed=5)
