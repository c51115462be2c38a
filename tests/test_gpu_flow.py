"""GPU parity of the Poisson solve, the momentum RHS and full RK3 steps
against the oracle (FP64, reference-faithful single projection)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NC = (32, 24, 16)
LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture()
def pair(po):
    from chapsim2_amd.domain import Domain
    from chapsim2_amd.initial import channel_fields

    dm = Domain(NC, LX, -1.0, 1.0, LZ, **CFG)
    of = po.Flow(list(NC), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **CFG)
    fields = channel_fields(NC, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=12345)
    dm.set_state(*fields)
    of.set_state(*fields)
    yield dm, of
    dm.close()


def test_poisson_vs_oracle(pair):
    from chapsim2_amd.poisson import solve_fft_poisson

    dm, of = pair
    rng = np.random.default_rng(5)
    rhs = np.asfortranarray(rng.standard_normal(NC))
    ref = of.poisson(rhs)
    out = np.array(rhs, order="F")
    solve_fft_poisson(out, dm)
    assert rel(out, ref) < 1e-12


@pytest.mark.parametrize("nc", [(128, 8, 64), (256, 8, 128), (512, 8, 256), (1024, 8, 512)])
def test_poisson_transform_lengths(po, nc):
    """FISHPACK-form Poisson solve where nx/2 and nz take each length of the
    two-pass register FFT kernels (64, 128, 256, 512) and the x forward /
    inverse real transforms; against the oracle's FFTPACK-form solve."""
    from chapsim2_amd.domain import Domain
    from chapsim2_amd.poisson import solve_fft_poisson

    dm = Domain(nc, LX, -1.0, 1.0, LZ, **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **CFG)
    try:
        rng = np.random.default_rng(7)
        rhs = np.asfortranarray(rng.standard_normal(nc))
        ref = of.poisson(rhs)
        out = np.array(rhs, order="F")
        solve_fft_poisson(out, dm)
        assert rel(out, ref) < 1e-12, rel(out, ref)
        # the launch shapes of these transforms were chosen by timing them
        import ctypes as C

        buf = C.create_string_buffer(1 << 16)
        assert dm.lib.chs_tune_report(buf, len(buf)) > 0
        report = buf.value.decode()
        for key in (f"fft_x+/f64/{nc[0] // 2}/", f"fft_x-/f64/{nc[0] // 2}/", f"fft_s+/f64/{nc[2]}/"):
            assert key in report, (key, report)
    finally:
        dm.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_poisson_mixed_radix_lengths_vs_oracle(po, variant):
    """FISHPACK-form Poisson solve where nx/2 and nz take the two-pass
    lengths with factors 3 and 5 (36, 40, 50, 72, 80, 90, 96, 100, 144, 180,
    192: register DFTs of 5, 6, 9, 10, 12 and 15 points as direct sums), each
    transform forced to one launch shape through chs_tune_force (0, 1, 2:
    the two-pass kernels with 4, 8, 16 lines per workgroup; 3: the general
    Stockham kernel), against the oracle's FFTPACK-form solve."""
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import Domain
    from chapsim2_amd.poisson import solve_fft_poisson

    lib = _lib.load()
    _lib.check(lib.chs_tune_force(b"fft_", variant))
    try:
        for nc in ((144, 8, 180), (72, 8, 192), (200, 8, 90), (80, 8, 100), (100, 8, 50), (192, 8, 40),
                   (180, 8, 144), (288, 8, 72), (160, 8, 36), (360, 8, 96), (384, 8, 80)):
            print(f"mixed-radix poisson variant {variant} nc={nc}", flush=True)
            dm = Domain(nc, LX, -1.0, 1.0, LZ, **CFG)
            try:
                of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **CFG)
                rng = np.random.default_rng(sum(nc) + variant)
                rhs = np.asfortranarray(rng.standard_normal(nc))
                ref = of.poisson(rhs)
                out = np.array(rhs, order="F")
                solve_fft_poisson(out, dm)
                assert rel(out, ref) < 1e-12, (variant, nc, rel(out, ref))
            finally:
                dm.close()
    finally:
        _lib.check(lib.chs_tune_force(b"fft_", -1))


def test_momentum_rhs_vs_oracle(pair):
    dm, of = pair
    for isub in (0, 1):
        ref = of.momentum_rhs(isub)
        got = dm.momentum_rhs(isub)
        for g, r, name in zip(got, ref, ("mx", "my", "mz")):
            assert rel(g, r) < 1e-12, (isub, name, rel(g, r))


def test_steps_vs_oracle(pair):
    dm, of = pair
    # initial Euler step + projection (src/chapsim.f90:135), then 3 RK3 steps
    of.update_re(0)
    dm.update_re(0)
    of.substep(0)
    dm.substep(0)
    for name, g, r in zip(("qx", "qy", "qz", "pr"), dm.get_state(), of.get_state()):
        assert rel(g, r) < 1e-12, ("init", name, rel(g, r))
    for it in range(1, 4):
        of.step(it)
        dm.run(it - 1, 1)
        for name, g, r in zip(("qx", "qy", "qz", "pr"), dm.get_state(), of.get_state()):
            assert rel(g, r) < 1e-12, (it, name, rel(g, r))
    np.testing.assert_allclose(dm.mass_check(), of.mass_check(), rtol=1e-6, atol=1e-13)


@pytest.mark.parametrize("nc,iacc", [((32, 24, 16), 4), ((48, 40, 24), 2), ((40, 32, 30), 1),
                                     ((32, 25, 16), 4), ((32, 144, 16), 4), ((128, 16, 64), 4)])
def test_exact_projection_steps_vs_oracle(po, nc, iacc):
    """Exact projection (inverse of the composed D*G): GPU eigen-basis solve
    against the oracle's per-mode dense LU; after every step max|div u| is at
    round-off (the reference's single 2nd-order projection grows it instead).
    Even ny takes the mirror-symmetric even/odd path (ny = 144: two row
    tiles per half), odd ny the general one.

    Tolerances: velocity 1e-13 relative (observed 1e-15 .. 3e-15 over the
    five steps); pressure 5e-11 (observed 2e-12 .. 1.0e-11, not growing):
    the pressure increment solves the nearly singular composed operator,
    whose smallest non-zero eigenvalues amplify round-off by their condition
    number, and the two sides reach it by different algorithms (eigen-basis
    products here, per-mode LU in the oracle)."""
    from chapsim2_amd.domain import PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    cfg = dict(CFG, iacc=iacc)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **cfg)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, poisson=2, projection=1, **cfg)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=99)
    try:
        dm.set_state(*fields)
        of.set_state(*fields)
        of.update_re(0)
        dm.update_re(0)
        of.substep(0)
        dm.substep(0)
        for it in range(1, 6):
            of.step(it)
            dm.run(it - 1, 1)
            for name, g, r in zip(("qx", "qy", "qz", "pr"), dm.get_state(), of.get_state()):
                assert rel(g, r) < (5e-11 if name == "pr" else 1e-13), (it, name, rel(g, r))
            div = dm.mass_check()
            assert np.all(div < 1e-12), (it, div)
    finally:
        dm.close()


def test_mixed_precision_steps_vs_fp64_oracle(po):
    """Mixed form (FP32 momentum right-hand side, FP64 state update,
    divergence, exact projection): after 10 RK3 steps the velocity stays
    within 1e-6 (relative) of the FP64 oracle and max|div u| at round-off."""
    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (48, 40, 32)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, precision=PREC_MIXED, **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=31)
    try:
        dm.set_state(*fields)
        of.set_state(*fields)
        of.update_re(0)
        dm.update_re(0)
        of.substep(0)
        dm.substep(0)
        for it in range(1, 11):
            of.step(it)
            dm.run(it - 1, 1)
        g, r = dm.get_state(), of.get_state()
        for name, a, b in zip(("qx", "qy", "qz"), g[:3], r[:3]):
            err = rel(a, b)
            assert err < 1e-6, (name, err)
            assert err > 1e-12  # the right-hand side really ran in FP32
        assert np.all(dm.mass_check() < 1e-12)
    finally:
        dm.close()


@pytest.mark.parametrize("nc,mixed", [((32, 24, 16), True), ((40, 300, 24), False), ((40, 300, 24), True)])
def test_momentum_rhs_paired_operators(po, nc, mixed):
    """The momentum right-hand side applies the interpolation and the first
    derivative of each velocity component along each direction in one pass
    (paired line kernels: y and z lines of 16-row chunks in both precisions
    and of 32-row chunks in FP32 (ny = 300), x lines staged 64 at a time
    where both plans fit the LDS; other pairs run one by one).  Against the
    FP64 oracle: FP64 within 1e-12, the mixed form's FP32 right-hand side
    within 1e-5 of the largest value."""
    from chapsim2_amd.domain import PREC_FP64, PREC_MIXED, Domain
    from chapsim2_amd.initial import channel_fields

    dm = Domain(nc, LX, -1.0, 1.0, LZ, precision=PREC_MIXED if mixed else PREC_FP64, **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **CFG)
    try:
        fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=515)
        dm.set_state(*fields)
        of.set_state(*fields)
        for isub in (0, 1):
            for g, r, name in zip(dm.momentum_rhs(isub), of.momentum_rhs(isub), ("mx", "my", "mz")):
                err = rel(g, r)
                print(f"{nc} mixed={mixed} isub {isub} {name}: {err:.2e}")
                assert err < (1e-5 if mixed else 1e-12), (isub, name, err)
                if mixed:
                    assert err > 1e-12  # the right-hand side really ran in FP32
    finally:
        dm.close()


def test_restart_resume_matches_continuous_run(po, tmp_path):
    """Checkpoint and resume through the reference's restart files: the flow
    written after two RK3 steps (chapsim2_amd.restart, the format of
    src/io_restart.f90:55-130) and loaded into a fresh solver continues
    bit-identically to an uninterrupted run (the first sub-step of an RK3
    step does not use the previous right-hand side, so q and p are the whole
    state between steps)."""
    from chapsim2_amd import restart
    from chapsim2_amd.domain import Domain
    from chapsim2_amd.initial import channel_fields

    fields = channel_fields(NC, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=808)
    a = Domain(NC, LX, -1.0, 1.0, LZ, **CFG)
    b = Domain(NC, LX, -1.0, 1.0, LZ, **CFG)
    try:
        a.set_state(*fields)
        a.update_re(0)
        a.substep(0)
        a.run(0, 2)
        restart.write_flow(a, 2, directory=str(tmp_path))
        a.run(2, 2)
        restart.read_flow(b, 2, directory=str(tmp_path))
        b.run(2, 2)
        for name, x, y in zip(restart.FLOW_FIELDS, a.get_state(), b.get_state()):
            assert np.array_equal(x, y), (name, np.abs(x - y).max())
    finally:
        a.close()
        b.close()


def test_cfl_vs_oracle(pair):
    """chs_cfl (convective CFL from the P2C-interpolated velocity, diffusion
    number) against the oracle's Check_cfl_convection / Check_cfl_diffusion
    (src/tools_solver.f90:220-375) on the noisy stretched channel, before and
    after two RK3 steps."""
    dm, of = pair
    for step in range(2):
        if step:
            of.update_re(0)
            dm.update_re(0)
            of.substep(0)
            dm.substep(0)
            of.step(1)
            dm.run(0, 1)
        g, r = dm.cfl(), of.cfl()
        assert abs(g[0] - r[0]) <= 1e-13 * r[0], (step, g, r)
        assert abs(g[1] - r[1]) <= 1e-14 * r[1], (step, g, r)
        assert r[0] > 0


def test_mixed_state_copies_follow_set_state():
    """Mixed form: a sub-step leaves the FP32 copies of the new velocity and
    pressure (stored by the corrections and the pressure update) for the
    next right-hand side, and set_state makes them stale.  A solver stepped
    twice, then given other fields and stepped again, must land bit for bit
    where a fresh solver given those fields and stepped once lands; and the
    momentum right-hand side from the copies a step left must equal the one
    from copies converted afresh."""
    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (32, 24, 16)
    kw = dict(projection=PROJ_EXACT, precision=PREC_MIXED, **CFG)
    a = Domain(nc, LX, -1.0, 1.0, LZ, **kw)
    b = Domain(nc, LX, -1.0, 1.0, LZ, **kw)
    try:
        f1 = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=71)
        f2 = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=72)
        a.set_state(*f1)
        a.substep(0)
        a.run(0, 2)
        a.set_state(*f2)
        b.set_state(*f2)
        a.run(2, 1)
        b.run(2, 1)
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y)
        b.run(3, 1)
        a.run(3, 1)
        b.set_state(*b.get_state())  # same state, copies converted afresh
        for x, y in zip(a.momentum_rhs(1), b.momentum_rhs(1)):
            assert np.array_equal(x, y)
    finally:
        a.close()
        b.close()


def test_mass_check_windows_closed_form_on_gpu():
    """chs_mass_check on qx = cos(2 pi x / LX) alone against the closed-form
    maxima of the three x windows (the construction that pins the oracle's
    windows in test_oracle_diagnostics)."""
    from chapsim2_amd.domain import Domain
    from test_oracle_diagnostics import LX as LXD, LZ as LZD, _window_reference

    nc = (48, 8, 8)
    dm = Domain(nc, LXD, -1.0, 1.0, LZD, istret=0, rstret=0.0, dt=1e-3, ren=100.0, idriven=0)
    try:
        qx, ref, _ = _window_reference(nc)
        dm.set_state(qx, np.zeros(dm.shapes["qy"], order="F"), np.zeros(dm.shapes["qz"], order="F"),
                     np.zeros(dm.shapes["pr"], order="F"))
        out = dm.mass_check()
        assert np.allclose(out, ref, rtol=1e-12, atol=0), (out, ref)
    finally:
        dm.close()


def test_velocity_range_and_nonfinite_detection(pair):
    """chs_velocity_range (the per-step Find_max_min_3d of qx, qy, qz) is
    exact: the max and min numpy takes of the state, before and after an RK3
    step.  A NaN or an infinity planted in the state makes it return
    CHS_ENAN with the range of the rest (NaN skipped, an infinity taken as
    the extreme it is), and the Python wrapper raise."""
    from chapsim2_amd import _lib

    dm, _ = pair

    def expect(qx, qy, qz):
        return [np.nanmax(qx), np.nanmin(qx), np.nanmax(qy), np.nanmin(qy), np.nanmax(qz), np.nanmin(qz)]

    qx, qy, qz, pr = dm.get_state()
    assert np.array_equal(dm.velocity_range(), expect(qx, qy, qz))
    dm.run(1, 1)
    s = dm.get_state()
    assert np.array_equal(dm.velocity_range(), expect(*s[:3]))
    qy2 = np.array(qy, order="F")
    qy2[3, 5, 7] = np.nan
    qz2 = np.array(qz, order="F")
    qz2[1, 2, 3] = -np.inf
    dm.set_state(qx, qy2, qz2, pr)
    out, rc = dm.velocity_range(check=False)
    assert rc == _lib.CHS_ENAN
    assert np.array_equal(out, expect(qx, qy2, qz2)) and out[5] == -np.inf
    with pytest.raises(_lib.ChsError):
        dm.velocity_range()
    dm.set_state(qx, qy, qz, pr)
    assert np.array_equal(dm.velocity_range(), expect(qx, qy, qz))


def test_bulk_monitor_vs_oracle(pair):
    """chs_bulk (bulk mean kinetic energy of the cell-centred velocity and
    bulk qx, the values of write_monitor_bulk) against the oracle on the
    noisy stretched channel."""
    dm, of = pair
    g, r = dm.bulk(), of.bulk()
    assert np.all(np.abs(g - r) <= 1e-13 * np.abs(r)), (g, r)
    assert abs(r[1] - 1.0) < 0.1  # the initial field carries a unit bulk velocity
