"""Many solves on one GPU at once (mctdhb_amd.batch): every instance's
result is the same alone and in the mix (bit for bit: instances share no
buffers and every kernel is deterministic), and the energy read-out follows
the oracle's total_energy (reference integrator_monitor.c:56-88)."""
import numpy as np
import pytest

import oracle
from mctdhb_amd import batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _draws():
    # small spaces (FP64 one-workgroup step) and two above 8192 (mixed form)
    small = batch.draw_ground_states(5, seed=3, max_dim=3000)
    big = [d for d in batch.draw_ground_states(40, seed=5, max_dim=30000) if d.dim > 8192][:2]
    assert len(big) == 2
    return small + big


def _state(inst):
    import torch

    torch.cuda.synchronize()
    return {k: inst.t[k].cpu().numpy().copy() for k in ("orb", "coef", "r1", "rinv")}


def test_batch_same_alone_and_in_the_mix(amd):
    import torch

    draws = _draws()
    print("batch draws", [(d.npar, d.norb, d.ngrid, d.dim) for d in draws], flush=True)
    alone = []
    for d in draws:
        inst = batch.Instance(d)
        for _ in range(12):
            inst.step()
        alone.append(_state(inst))
        inst.close()
    insts = [batch.Instance(d) for d in draws]
    b = batch.Batch(insts, nstreams=3)
    b.run(12)
    b.synchronize()
    torch.cuda.synchronize()
    for d, inst, ref in zip(draws, insts, alone):
        got = _state(inst)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (d, k)
    for inst in insts:
        inst.close()


def test_batch_energy_readout_and_descent(amd):
    d = batch.draw_ground_states(1, seed=9, max_dim=2000)[0]
    print("energy draw", d, flush=True)
    inst = batch.Instance(d)
    for _ in range(5):
        inst.step()
    import torch

    torch.cuda.synchronize()
    e5 = inst.energy_per_particle()
    M = d.norb
    h = {k: inst.t[k].cpu().numpy() for k in ("r1", "r2", "hob", "hint")}
    ref = oracle.total_energy(M, h["r1"], h["r2"], h["hob"], h["hint"]).real / d.npar
    assert abs(e5 - ref) <= 1e-12 * max(1.0, abs(ref))
    for _ in range(200):
        inst.step()
    torch.cuda.synchronize()
    e205 = inst.energy_per_particle()
    assert e205 < e5  # imaginary time lowers the energy
    occ = inst.occupations()
    assert abs(occ.sum() - 1.0) < 1e-10 and np.all(occ > -1e-12)
    inst.close()
