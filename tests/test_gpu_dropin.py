"""The reference-named whole-step entry points (get_mctdhb_struct,
sync_density_matrices, sync_orbital_matrices, mctdhb_propagate_step,
total_energy) on host structs against the oracle's step, through the C ABI."""
import numpy as np
import pytest

import gpu_util as gu
from mctdhb_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("imagtime,orb_der,rk", [(True, 301, 4), (False, 302, 4), (True, 301, 5), (False, 301, 2)])
def test_reference_step_matches_oracle(orc, imagtime, orb_der, rk):
    print(f"drop-in step imag={imagtime} der={orb_der} rk={rk}", flush=True)
    from mctdhb_amd import dropin

    npar, norb, grid, lm, dt, g = 5, 3, 129, 5, 0.01, 1.0
    x = np.linspace(-8, 8, grid)
    dx = x[1] - x[0]
    m = dropin.MCTDHB(npar, norb, grid, x[0], x[-1], dt, 1.0, lambda t, xs: 0.5 * xs ** 2, lambda t: g,
                      imagtime=imagtime, orb_der=orb_der, rk_order=rk, lanczos_iter=lm)
    orb = np.ascontiguousarray(inputs.hermite_gauss(norb, x, seed=1, noise=1e-3).ravel())
    orc.lib().orc_orthonormalize(grid, dx, norb, orb)
    S = orc.Space(npar, norb)
    coef = inputs.coefficients(S.dim, 1)
    m.set("orbitals", orb)
    m.set("coef", coef)
    m.sync()

    eq = orc.make_orbeq(norb, grid, x[0], x[-1], dt, g, 0.5 * x ** 2, imagtime=imagtime, rk_order=rk)
    r1, r2, rinv = S.sync_density(coef)
    r1 = np.ascontiguousarray(r1.ravel())
    rinv = np.ascontiguousarray(rinv.ravel())
    hob = np.zeros(norb * norb, complex)
    hint = np.zeros(norb ** 4, complex)
    orc.lib().orc_orbital_hob(eq, orb, hob)
    orc.lib().orc_orbital_hint(eq, orb, hint)
    assert gu.rel_err(m.get("ob_denmat").ravel(), r1) < 1e-12
    assert gu.rel_err(m.get("hob").ravel(), hob) < 1e-12
    assert gu.rel_err(m.get("hint"), hint) < 1e-12

    lvec = np.zeros(lm * S.dim, complex)
    for _ in range(4):
        orc.lib().orc_propagate_step(S._p, eq, int(imagtime), int(orb_der != 302), lm, orb, coef, hob, hint, r1,
                                     r2, rinv, lvec)
        m.step()
    got = {"orbitals": m.get("orbitals").ravel(), "coef": m.get("coef"), "hob": m.get("hob").ravel(),
           "hint": m.get("hint"), "ob_denmat": m.get("ob_denmat").ravel(), "tb_denmat": m.get("tb_denmat"),
           "inv_ob_denmat": m.get("inv_ob_denmat").ravel()}
    ref = {"orbitals": orb, "coef": coef, "hob": hob, "hint": hint, "ob_denmat": r1, "tb_denmat": r2,
           "inv_ob_denmat": rinv}
    for k in ref:
        assert gu.rel_err(got[k], ref[k]) < 1e-9, k
    e = m.energy()
    e_ref = orc.total_energy(norb, r1, r2, hob, hint)
    assert abs(e - e_ref) < 1e-9 * abs(e_ref)
    m.close()
