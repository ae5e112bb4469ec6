"""Packed steps (mctdhb_pack_*): one kernel launch per stage advances many
small solves by one whole step (reference driver.c:159-202 per solve).
Every stage runs the block code of the single-solve kernels, so each solve's
state after packed steps equals, bit for bit, its state after the same
number of its own step graph replays."""
import numpy as np
import pytest

from mctdhb_amd import batch

pytestmark = pytest.mark.gpu

KEYS = ("orb", "coef", "hob", "hint", "r1", "r2", "rinv")


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _state(inst):
    import torch

    torch.cuda.synchronize()
    return {k: inst.t[k].cpu().numpy().copy() for k in KEYS}


def _draws():
    # M = 2..6 over the three grid lengths, with spaces on both sides of the
    # one-workgroup SIL kernel's staging limit, plus M = 7 (the RK sub-step
    # with its tables formed by a kernel of their own)
    ds = batch.draw_ground_states(14, seed=21, max_dim=8192)
    ds.append(batch.Draw(npar=6, norb=7, ngrid=256, a=1.0, b=2.0, g=0.5, seed=77))
    ds.append(batch.Draw(npar=40, norb=3, ngrid=512, a=1.2, b=1.5, g=2.0, seed=78))  # NC 861
    return ds


def test_pack_matches_step_graphs(amd):
    draws = _draws()
    print("pack draws", [(d.npar, d.norb, d.ngrid, d.dim) for d in draws], flush=True)
    alone = []
    for d in draws:
        inst = batch.Instance(d)
        for _ in range(6):
            inst.step()
        alone.append(_state(inst))
        inst.close()
    insts = [batch.Instance(d) for d in draws]
    b = batch.Batch(insts, nstreams=2, packed=True, pack_streams=4)
    assert b.packed_count == len(draws)
    assert b.pack.groups >= 5
    print("pack groups", b.pack.groups, flush=True)
    b.run(6)
    b.synchronize()
    for d, inst, ref in zip(draws, insts, alone):
        got = _state(inst)
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (d, k)
    b.close()
    for inst in insts:
        inst.close()


def test_pack_beside_graphs_and_real_time(amd):
    """A batch with instances above the one-workgroup size (their graphs
    beside the pack), then a real-time quench of packed solves."""
    import torch

    small = batch.draw_ground_states(4, seed=31, max_dim=3000)
    big = [d for d in batch.draw_ground_states(40, seed=5, max_dim=30000) if d.dim > 8192][:1]
    draws = small + big
    print("mixed batch draws", [(d.npar, d.norb, d.ngrid, d.dim) for d in draws], flush=True)
    alone = []
    for d in draws:
        inst = batch.Instance(d, quench_g=2.0 * d.g)
        for _ in range(4):
            inst.step()
        for _ in range(3):
            inst.step_real()
        alone.append(_state(inst))
        inst.close()
    insts = [batch.Instance(d, quench_g=2.0 * d.g) for d in draws]
    b = batch.Batch(insts, nstreams=2, packed=True)
    assert b.packed_count == len(small)
    assert not amd.pack_eligible(insts[-1].space, insts[-1].sil, insts[-1].orbital)
    b.run(4)
    b.synchronize()
    b.close()
    # real time: the quench orbital workspaces in a pack of their own
    solves = [(i.space, i.sil, i.orbital_real, i.t["orb"], i.t["coef"], i.t["hob"], i.t["hint"], i.t["r1"],
               i.t["r2"], i.t["rinv"]) for i in insts[:-1]]
    pr = amd.Pack(solves)
    s = torch.cuda.Stream()
    pr.step([s], 3)
    for _ in range(3):
        insts[-1].step_real(s)
    s.synchronize()
    pr.close()
    for d, inst, ref in zip(draws, insts, alone):
        got = _state(inst)
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (d, k)
    for inst in insts:
        inst.close()


def test_pack_refuses_what_it_does_not_cover(amd):
    d_big = [d for d in batch.draw_ground_states(40, seed=5, max_dim=30000) if d.dim > 8192][0]
    d_odd = batch.Draw(npar=8, norb=3, ngrid=200, a=1.0, b=2.0, g=0.3, seed=5)  # FFT length 200
    for d in (d_big, d_odd):
        inst = batch.Instance(d)
        assert not amd.pack_eligible(inst.space, inst.sil, inst.orbital)
        t = inst.t
        with pytest.raises(amd.MctdhbError):
            amd.Pack([(inst.space, inst.sil, inst.orbital, t["orb"], t["coef"], t["hob"], t["hint"], t["r1"],
                       t["r2"], t["rinv"])])
        inst.close()
