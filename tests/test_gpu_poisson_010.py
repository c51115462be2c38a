"""GPU parity of the channel path of the 3-D FFT Poisson library
(poisson_010, src/poisson_1stderivcomp.f90:1457-1816: x and z periodic, y
walls on a uniform mesh) against the oracle's literal restatement, which is
pinned by the mirrored triply periodic solve (test_oracle_fft_poisson.py),
and RK3 steps of a uniform channel that project with it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=0, dt=5e-4, ren=2800.0, idriven=1)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ny / 2 <= 64 takes the FP64 matrix-core mixing, ny / 2 > 64 the two-pass
# even/odd kernels; odd nz and the other accuracies are covered
@pytest.mark.parametrize("nc,iacc", [((32, 24, 16), 4), ((48, 40, 30), 2), ((64, 32, 45), 1),
                                     ((32, 160, 16), 4), ((64, 130, 48), 3), ((256, 128, 256), 4)])
def test_poisson_010_vs_oracle(po, nc, iacc):
    """FP64 solve: within 1e-12 of the oracle (relative to the largest value)."""
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain
    from chapsim2_amd.poisson import solve_fft_poisson

    print(f"poisson_010 nc={nc} iacc={iacc}", flush=True)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, iacc=iacc, poisson=FFT_2DECOMP_3DFFT, **CFG)
    try:
        of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=iacc, poisson=3, **CFG)
        rng = np.random.default_rng(sum(nc) + iacc)
        rhs = np.asfortranarray(rng.standard_normal(nc))
        ref = of.poisson(rhs)
        out = np.array(rhs, order="F")
        solve_fft_poisson(out, dm)
        err = rel(out, ref)
        print(f"  rel {err:.2e}", flush=True)
        assert err < 1e-12, err
    finally:
        dm.close()


@pytest.mark.parametrize("nc", [(64, 48, 40), (32, 160, 16)])
def test_poisson_010_fp32_vs_oracle(po, nc):
    """FP32 solve (chs_poisson_f32: the one-pass FP32 even/odd mixing for
    ny / 2 <= 64, the two-pass one above) against the oracle's FP64 solve:
    FP32 rounding, above FP64 round-off."""
    import ctypes as C

    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain

    dm = Domain(nc, LX, -1.0, 1.0, LZ, iacc=4, poisson=FFT_2DECOMP_3DFFT, **CFG)
    try:
        of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=3, **CFG)
        rng = np.random.default_rng(77)
        rhs = rng.standard_normal(nc)
        rhs -= rhs.mean()
        rhs = np.asfortranarray(rhs)
        ref = of.poisson(rhs)
        t = torch.from_numpy(np.ascontiguousarray(rhs.T)).float().cuda()
        _lib.check(dm.lib.chs_poisson_f32(dm.h, C.c_void_p(t.data_ptr())))
        got = t.double().cpu().numpy().T
        err = rel(got, ref)
        print(f"f32 poisson_010: {err:.2e}", flush=True)
        assert 1e-9 < err < 1e-5, err
    finally:
        dm.close()


@pytest.mark.parametrize("nc,iacc", [((32, 24, 16), 4), ((40, 32, 30), 2)])
def test_uniform_channel_steps_on_poisson_010_vs_oracle(po, nc, iacc):
    """A uniform channel (constant flux) whose sub-steps project with
    poisson_010: the initial sub-step and 3 RK3 steps, velocity within 1e-12
    of the oracle flow on the same path. The pressure (mean removed) is held
    to 1e-12 or ten times the oracle's own response to a one-ulp change of
    the initial velocity, whichever is larger: each increment solves
    div(u*) / (alpha dt), a small difference of O(1) velocities, so its
    relative rounding exceeds the velocity's."""
    from tolerances import rel_mean_removed, ulp_perturbed

    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain
    from chapsim2_amd.initial import channel_fields

    dm = Domain(nc, LX, -1.0, 1.0, LZ, iacc=iacc, poisson=FFT_2DECOMP_3DFFT, **CFG)
    try:
        of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=iacc, poisson=3, **CFG)
        op_ = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=iacc, poisson=3, **CFG)
        fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 0, 0.0, noise=0.1, seed=4321)
        dm.set_state(*fields)
        of.set_state(*fields)
        op_.set_state(*ulp_perturbed(fields))
        for x in (dm, of, op_):
            x.update_re(0)
            x.substep(0)
        for it in range(1, 4):
            of.step(it)
            op_.step(it)
            dm.run(it - 1, 1)
            got, ref = dm.get_state(), of.get_state()
            errs = {k: rel(g, r) for k, g, r in zip(("qx", "qy", "qz"), got[:3], ref[:3])}
            errs["pr"] = rel_mean_removed(got[3], ref[3])
            ptol = max(1e-12, 10.0 * rel_mean_removed(op_.get_state()[3], ref[3]))
            print(f"poisson_010 channel {nc} step {it}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items())
                  + f" (pressure tolerance {ptol:.1e})", flush=True)
            assert max(errs[k] for k in ("qx", "qy", "qz")) < 1e-12, (it, errs)
            assert errs["pr"] < ptol, (it, errs, ptol)
        np.testing.assert_allclose(dm.mass_check(), of.mass_check(), rtol=1e-6, atol=1e-13)
        print("max|div u| after 3 steps", dm.mass_check(), flush=True)
    finally:
        dm.close()


def test_uniform_channel_mixed_on_poisson_010_vs_fp64_oracle(po):
    """The mixed form (FP32 right-hand side, FP64 state and projection) on
    the poisson_010 path: 10 RK3 steps within 1e-6 of the FP64 oracle in
    velocity. The divergence this projection leaves (a few 1e-9 with CP6:
    the cosine solve inverts the periodic modified wavenumbers, not the
    wall rows of the compact operators) matches the oracle's to 1e-10
    absolute, the mixed form's own FP32 rounding moving it by ~1e-11."""
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, PREC_MIXED, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (32, 24, 16)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, iacc=4, poisson=FFT_2DECOMP_3DFFT, precision=PREC_MIXED, **CFG)
    try:
        of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=3, **CFG)
        fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 0, 0.0, noise=0.1, seed=99)
        dm.set_state(*fields)
        of.set_state(*fields)
        for x in (dm, of):
            x.update_re(0)
            x.substep(0)
        dm.run(0, 10)
        for it in range(1, 11):
            of.step(it)
        errs = {k: rel(g, r) for k, g, r in zip(("qx", "qy", "qz"), dm.get_state()[:3], of.get_state()[:3])}
        print("poisson_010 mixed 10 steps:", " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
        assert max(errs.values()) < 1e-6, errs
        assert 1e-12 < max(errs.values())  # the right-hand side really was FP32
        dg, dr = dm.mass_check(), of.mass_check()
        print("max|div u|", dg, "oracle", dr, flush=True)
        assert np.abs(dg - dr).max() < 1e-10, (dg, dr)
    finally:
        dm.close()


def test_poisson_010_rejections(po):
    """What the channel 3-D FFT path does not cover is refused at creation:
    a stretched y mesh (the reference switches to its 3FMD mapping with a
    pentadiagonal y solve there), an odd ny (its y reorder pairs the rows),
    and the exact projection on this set-up."""
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, PROJ_EXACT, Domain

    bad = [dict(nc=(32, 24, 16), istret=2, rstret=0.075), dict(nc=(32, 25, 16), istret=0),
           dict(nc=(32, 24, 16), istret=0, projection=PROJ_EXACT)]
    for b in bad:
        kw = dict(b)
        nc = kw.pop("nc")
        with pytest.raises(_lib.ChsError) as ei:
            Domain(nc, LX, -1.0, 1.0, LZ, iacc=4, poisson=FFT_2DECOMP_3DFFT, dt=5e-4, ren=2800.0, **kw)
        assert ei.value.code == _lib.CHS_EUNSUP, (b, ei.value)
