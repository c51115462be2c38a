"""GPU parity of the variable-property (heated channel) path against the
oracle: the energy right-hand side, the thermo sub-step sequence (energy
equation, property refresh, variable-density / variable-viscosity momentum,
projection of g = rho u, q = g / rho), over several RK3 steps, FP64 and mixed;
and the surrogate property mode against the table mode.

The enthalpy field is non-uniform from the start (h between 0.77 and 1.18 on
the heated channel's property table) so that every property-dependent term
is exercised.  Pressure is compared as in the long-horizon tests: mean
removed, held to 4 kappa(L_y) eps per projection.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
EPS64 = 2.0 ** -52


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _table():
    from chapsim2_amd.initial import HEATED_H_RANGE, heated_channel_table

    return heated_channel_table(4096, *HEATED_H_RANGE)


def _initial(po, nc, seed):
    """noisy channel velocity and a non-uniform rho*h: h = 0.8 + 0.35 y^2
    + 0.03 noise at the cell centres"""
    from chapsim2_amd.initial import channel_fields

    tab = _table()
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=seed)
    dom = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    rng = np.random.default_rng(seed + 1)
    h = 0.8 + 0.35 * dom.yc[None, :, None] ** 2 + 0.03 * rng.uniform(-1, 1, nc)
    rhoh = np.asfortranarray(np.interp(h, tab[:, 1], tab[:, 3] * tab[:, 1]))
    tw = float(np.interp(1.2, tab[:, 1], tab[:, 2]))
    t0 = float(np.interp(0.8, tab[:, 1], tab[:, 2]))
    return tab, fields, rhoh, (tw, tw), t0


def _pair(po, nc, seed, precision=0, mode=0, surrogate_steps=0):
    from chapsim2_amd.domain import PROJ_EXACT, Domain
    from chapsim2_amd.props import PropertyTable

    tab, fields, rhoh, tw, t0 = _initial(po, nc, seed)
    pt = PropertyTable(tab, -1.0)
    sur = None
    if mode == 1:
        from chapsim2_amd.mlp import Surrogate

        sur = Surrogate(pt, nsamples=1_000_000, seed=3)
        sur.train(surrogate_steps, batch=4096, lr=2e-3)
        pt.set_surrogate(sur)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, precision=precision, **CFG)
    dm.set_state(*fields)
    dm.thermo_enable(pt, tw, 1.0, t0, mode)
    dm.thermo_set_rhoh(rhoh)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    of.set_state(*fields)
    opt = po.Props(tab, -1.0)
    of.thermo_enable(opt, tw, 1.0, t0)
    assert of.thermo_set_rhoh(rhoh) == 0
    return dm, of, (pt, sur)


def test_energy_rhs_vs_oracle(po):
    """Compute_energy_rhs with the fractional step (sub-step 1) on the
    non-uniform heated channel: convection of rho u h, conduction with the
    interpolated conductivity, the isothermal wall values."""
    dm, of, keep = _pair(po, (32, 24, 16), 5)
    try:
        g = dm.energy_rhs(1)
        r = of.energy_rhs(1)
        err = rel(g, r)
        print(f"energy rhs 32x24x16: rel {err:.2e} (max |rhs| {np.abs(r).max():.3e})", flush=True)
        assert err < 1e-12
        # the property refresh of the initial state agrees too
        tg, tr = dm.thermo_get(), of.thermo_get()
        for k in ("h", "t", "k", "d", "m", "gx", "gy", "gz"):
            assert rel(tg[k], tr[k]) < 1e-13, (k, rel(tg[k], tr[k]))
    finally:
        dm.close()


def _compare(tag, dm, of, n_proj, vtol):
    from tolerances import ly_kappa

    got, ref = dm.get_state(), of.get_state()
    tg, tr = dm.thermo_get(), of.thermo_get()
    errs = {k: rel(g, r) for k, g, r in zip(("qx", "qy", "qz"), got[:3], ref[:3])}
    errs["pr"] = rel(got[3] - got[3].mean(), ref[3] - ref[3].mean())
    for k in ("rhoh", "h", "t", "d", "m", "gx", "gy", "gz"):
        errs[k] = rel(tg[k], tr[k])
    ptol = 4.0 * ly_kappa(of.dom) * EPS64 * n_proj
    print(tag, " ".join(f"{k} {v:.2e}" for k, v in errs.items()), f"(pressure tolerance {ptol:.1e})", flush=True)
    bad = [k for k, v in errs.items() if k != "pr" and not v < vtol]
    if not errs["pr"] < max(ptol, vtol if vtol > 1e-9 else 0.0):
        bad.append("pr")
    return bad


@pytest.mark.parametrize("nc", [(32, 24, 16), (48, 40, 32)])
def test_thermo_steps_vs_oracle(po, nc):
    """The initial sub-step and 5 RK3 steps of the heated channel (energy
    equation, property refresh, variable-property momentum, exact projection
    of g): every field within 1e-12 of the oracle, max|div g| at round-off."""
    dm, of, keep = _pair(po, nc, 9)
    try:
        for x in (dm, of):
            x.update_re(0)
            x.substep(0)
        dm.run(0, 5)
        for it in range(1, 6):
            of.step(it)
        bad = _compare(f"thermo {nc} 5 steps fp64:", dm, of, 16, 1e-12)
        div = dm.mass_check()
        print("max|div g|", div, flush=True)
        assert not bad, bad
        assert np.all(div < 1e-12)
    finally:
        dm.close()


def test_thermo_mixed_vs_fp64_oracle(po):
    """Mixed precision (FP32 momentum right-hand side with the variable-
    property terms, FP32 energy right-hand side; rho h, properties and
    projection FP64): 10 RK3 steps
    within 1e-6 of the FP64 oracle in velocity and in the thermal fields."""
    from chapsim2_amd.domain import PREC_MIXED

    dm, of, keep = _pair(po, (32, 24, 16), 13, precision=PREC_MIXED)
    try:
        for x in (dm, of):
            x.update_re(0)
            x.substep(0)
        dm.run(0, 10)
        for it in range(1, 11):
            of.step(it)
        bad = _compare("thermo 32x24x16 10 steps mixed:", dm, of, 31, 1e-6)
        bad = [b for b in bad if b != "pr"]
        assert not bad, bad
        assert np.all(dm.mass_check() < 1e-12)
    finally:
        dm.close()


def test_thermo_surrogate_mode_vs_table(po):
    """The surrogate (trained on the table's own interpolation) in place of
    the table for the properties from h: 10 RK3 steps against the table mode
    on the same state.  The properties start off by the surrogate's fit
    error over the enthalpies present (measured here) and the velocity
    q = g / rho by the density's.  The refresh takes h = rho*h / rho, so a
    density error moves h by the same relative amount and every property
    with it (through dp/dh); over the steps the differences feed back
    through rho, h and g (4-12 times the driving fit error after 10 steps
    when measured).  Each property is held to ten times the larger of its
    own fit error and the density's, the velocities (relative to the
    largest velocity component) to ten times the density's.  Reported per
    field."""
    from chapsim2_amd.domain import PROJ_EXACT, Domain

    dm, of, keep = _pair(po, (32, 24, 16), 17, mode=1, surrogate_steps=6000)
    try:
        pt, sur = keep
        tab, fields, rhoh, tw, t0 = _initial(po, (32, 24, 16), 17)
        ref = Domain((32, 24, 16), LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **CFG)
        ref.set_state(*fields)
        ref.thermo_enable(pt, tw, 1.0, t0, 0)
        ref.thermo_set_rhoh(rhoh)
        for x in (dm, ref):
            x.update_re(0)
            x.substep(0)
            x.run(0, 10)
        a, b = dm.thermo_get(), ref.thermo_get()
        hgrid = np.linspace(b["h"].min(), b["h"].max(), 2001)
        fit, tbl = sur(hgrid), pt.from_h(hgrid)
        fit_err = {k: rel(fit[k], tbl[k]) for k in ("t", "d", "m", "k")}
        qa, qb = dm.get_state(), ref.get_state()
        errs = {k: rel(a[k], b[k]) for k in ("t", "d", "m", "k")}
        # velocity differences relative to the velocity scale (the largest
        # component, qx): a density error changes q = g / rho by the same
        # relative amount of the local velocity, which for qy and qz is far
        # below their own small maxima
        vscale = max(np.abs(x).max() for x in qb[:3])
        errs.update({k: np.abs(x - y).max() / vscale for k, x, y in zip(("qx", "qy", "qz"), qa[:3], qb[:3])})
        print("surrogate fit error over the state's h:", " ".join(f"{k} {v:.2e}" for k, v in fit_err.items()),
              flush=True)
        print("surrogate vs table after 10 steps:", " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
        for k in ("t", "d", "m", "k"):
            drive = max(fit_err[k], fit_err["d"])
            assert errs[k] < 10.0 * drive + 1e-12, (k, errs[k], fit_err[k], fit_err["d"])
        for k in ("qx", "qy", "qz"):
            assert errs[k] < 10.0 * fit_err["d"] + 1e-12, (k, errs[k], fit_err["d"])
        assert max(fit_err.values()) < 5e-2
        assert np.all(dm.mass_check() < 1e-12)
        ref.close()
    finally:
        dm.close()


@pytest.mark.parametrize("matrix_cores,tol", [(0, 3e-5), (1, 1e-3)])
def test_thermo_surrogate_mixed_vs_fp64(po, matrix_cores, tol):
    """The surrogate mode in the mixed form, whose refresh kernel stores the
    FP32 copies of h, mu, kappa and T that the FP32 right-hand sides read,
    against the surrogate mode in FP64 with the same trained network: the
    two differ by the FP32 right-hand sides, h by ~1e-7 after 10 RK3 steps.
    With FP32 activations (the vector kernel) the network is smooth in h
    and the properties and velocities stay within 3e-5; with BF16
    activations (the matrix cores) a change of h that small can move an
    activation across a BF16 rounding boundary, a jump of up to ~1e-4 in a
    property (6e-5 measured), so the bound there is 1e-3.  Both lie above
    FP64 round-off; a swapped or missing FP32 copy puts the runs O(1)
    apart."""
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT, Domain

    dm, of, keep = _pair(po, (32, 24, 16), 19, precision=PREC_MIXED, mode=1, surrogate_steps=1500)
    try:
        pt, sur = keep
        _lib.check(_lib.load().chs_mlp_set_matrix_cores(sur.h, matrix_cores))
        tab, fields, rhoh, tw, t0 = _initial(po, (32, 24, 16), 19)
        dm.thermo_set_rhoh(rhoh)  # initial properties and g = rho q again, with this kernel
        ref = Domain((32, 24, 16), LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **CFG)
        ref.set_state(*fields)
        ref.thermo_enable(pt, tw, 1.0, t0, 1)
        ref.thermo_set_rhoh(rhoh)
        for x in (dm, ref):
            x.update_re(0)
            x.substep(0)
            x.run(0, 10)
        a, b = dm.thermo_get(), ref.thermo_get()
        errs = {k: rel(a[k], b[k]) for k in ("h", "t", "d", "m", "k")}
        errs.update({k: rel(x, y) for k, x, y in zip(("qx", "qy", "qz"), dm.get_state()[:3], ref.get_state()[:3])})
        print(f"surrogate mode (matrix cores {matrix_cores}), mixed vs FP64 after 10 steps:",
              " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
        assert max(errs.values()) < tol, errs
        assert 1e-12 < max(errs[k] for k in ("qx", "qy", "qz")), errs
        assert np.all(dm.mass_check() < 1e-12)
        ref.close()
    finally:
        _lib.check(_lib.load().chs_mlp_set_matrix_cores(keep[1].h, 1))
        dm.close()


def test_thermo_full_size_vs_oracle(po):
    """256x128x256 heated channel: the initial sub-step and one RK3 step
    against the oracle at the benchmark's size."""
    dm, of, keep = _pair(po, (256, 128, 256), 21)
    try:
        for x in (dm, of):
            x.update_re(0)
            x.substep(0)
        dm.run(0, 1)
        of.step(1)
        bad = _compare("thermo 256x128x256 1 step fp64:", dm, of, 4, 1e-12)
        assert not bad, bad
    finally:
        dm.close()
