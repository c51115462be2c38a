"""Triply periodic Taylor-Green vortex (config of src/initialisation.f90
:892-975 with the 3-D FFT Poisson path poisson_000): GPU against the
oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 2 * np.pi
CFG = dict(istret=0, rstret=0.0, dt=1e-3, ren=1600.0, idriven=0)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nc,iacc", [((32, 32, 32), 4), ((48, 24, 40), 4), ((32, 32, 32), 2), ((64, 64, 64), 4)])
def test_tgv_steps_vs_oracle(po, nc, iacc):
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain
    from chapsim2_amd.initial import tgv_fields
    from chapsim2_amd.poisson import solve_fft_poisson

    cfg = dict(CFG, iacc=iacc)
    dm = Domain(nc, L, -np.pi, np.pi, L, periodic_y=True, poisson=FFT_2DECOMP_3DFFT, **cfg)
    of = po.Flow(list(nc), L, -np.pi, np.pi, L, wall_y=False, poisson=3, **cfg)
    try:
        rng = np.random.default_rng(2)
        rhs = np.asfortranarray(rng.standard_normal(nc))
        ref = of.poisson(rhs)
        out = np.array(rhs, order="F")
        solve_fft_poisson(out, dm)
        assert rel(out, ref) < 1e-12
        fields = tgv_fields(nc, L, -np.pi, np.pi, L)
        dm.set_state(*fields)
        of.set_state(*fields)
        dm.update_re(0)
        of.update_re(0)
        dm.substep(0)
        of.substep(0)
        for it in range(1, 4):
            of.step(it)
            dm.run(it - 1, 1)
            for name, g, r in zip(("qx", "qy", "qz", "pr"), dm.get_state(), of.get_state()):
                assert rel(g, r) < 1e-12, (it, name, rel(g, r))
        assert np.all(dm.mass_check() < 1e-12)
    finally:
        dm.close()
