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


def _solve(amd, npar, norb, grid, bounds, rk_order, imagtime, seed):
    """One solve built from the low-level API, its initial state formed."""
    import math

    import torch

    from mctdhb_amd import inputs

    x = np.linspace(-8.0, 8.0, grid)
    pot = 0.5 * x ** 2
    space = amd.DeviceSpace(npar, norb, tables=False)
    sil = amd.DeviceSil(space, 5, amd.FP64, amd.INDEX_FLY)
    orbital = amd.DeviceOrbital(norb, grid, x[0], x[-1], 0.005, 0.7, pot, imagtime=imagtime, bounds=bounds,
                                rk_order=rk_order)
    orb = inputs.hermite_gauss(norb, x, seed=seed, noise=1e-3)
    if bounds == 1001:
        orb[:, -1] = orb[:, 0]
    dim = math.comb(npar + norb - 1, norb - 1)
    t = {"orb": torch.from_numpy(np.ascontiguousarray(orb.ravel())).cuda(),
         "coef": torch.from_numpy(inputs.coefficients(dim, seed)).cuda()}
    for k, n in (("hob", norb ** 2), ("hint", norb ** 4), ("r1", norb ** 2), ("r2", norb ** 4),
                 ("rinv", norb ** 2)):
        t[k] = torch.zeros(n, dtype=torch.complex128, device="cuda")
    orbital.orthonormalize(t["orb"])
    orbital.orbital_matrices(t["orb"], t["hob"], t["hint"])
    amd.density_sync(space, t["coef"], t["r1"], t["r2"], t["rinv"])
    torch.cuda.synchronize()
    return {"objs": (space, sil, orbital), "t": t}


def test_pack_rk_orders_bounds_and_real_time(amd):
    """Runge-Kutta orders 2 and 5 (2 and 6 stages) beside 4, zero and
    periodic bounds, imaginary and real time, in one pack."""
    import torch

    cases = [(8, 3, 257, 1000, 2, True), (6, 4, 129, 1001, 5, True), (10, 2, 257, 1000, 4, False),
             (5, 5, 513, 1001, 5, False), (7, 3, 129, 1001, 2, False)]
    print("pack cases", cases, flush=True)

    def tup(s):
        t = s["t"]
        return (*s["objs"], t["orb"], t["coef"], t["hob"], t["hint"], t["r1"], t["r2"], t["rinv"])

    ref = [_solve(amd, *c, seed=40 + i) for i, c in enumerate(cases)]
    graphs = [amd.StepGraph(*tup(s)) for s in ref]
    for _ in range(4):
        for g in graphs:
            g.launch()
    torch.cuda.synchronize()
    packed = [_solve(amd, *c, seed=40 + i) for i, c in enumerate(cases)]
    for s in packed:
        assert amd.pack_eligible(*s["objs"])
    pk = amd.Pack([tup(s) for s in packed])
    assert pk.groups == len(cases)  # every case differs in M, stage count or time direction
    st = torch.cuda.Stream()
    pk.step([st], 4)
    st.synchronize()
    for c, a, b in zip(cases, ref, packed):
        for k in KEYS:
            assert np.array_equal(a["t"][k].cpu().numpy(), b["t"][k].cpu().numpy()), (c, k)
    pk.close()
    for g in graphs:
        g.close()


def test_pack_refuses_shared_workspaces(amd):
    """Two entries over one space (its density partial sums) or one
    orbital workspace would race inside a packed launch."""
    d = batch.draw_ground_states(1, seed=3, max_dim=2000)[0]
    a, b = batch.Instance(d), batch.Instance(d)

    def tup(i, space=None, orbital=None):
        t = i.t
        return (space or i.space, i.sil, orbital or i.orbital, t["orb"], t["coef"], t["hob"], t["hint"], t["r1"],
                t["r2"], t["rinv"])

    for bad in ([tup(a), tup(b, space=a.space)], [tup(a), tup(b, orbital=a.orbital)], [tup(a), tup(a)]):
        with pytest.raises(amd.MctdhbError):
            amd.Pack(bad)
    amd.Pack([tup(a), tup(b)]).close()
    a.close()
    b.close()
