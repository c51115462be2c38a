"""GPU parity at the benchmark's full size, over long horizons and on every
batched-channel grid, against the CPU oracle.

- FP64 with the exact projection at 256x128x256 (the channel benchmark's own
  grid): the initial step and two RK3 steps.
- The channel workload over 200 RK3 steps on a reduced grid, FP64 and mixed.
- The Taylor-Green vortex on 64^3 for its 100 steps.
- The 16 uneven channels of the batched benchmark (FFT lengths with the
  3^2 and 5^2 Stockham stages), one eager and two graph-captured steps, FP64
  and mixed, every instance against the oracle.
- The exact projection on a grid with more y-mixing tiles than compute units
  (the persistent multi-tile loop and its next-tile prefetch).

Pressure: the exact projection's pressure increment solves the composed
operator L = D*G, whose wall-normal block L_y (D1_P2C after D1_C2P with the
wall faces reset) has a one-dimensional null space (the gauge) and, on its
range, the condition number kappa = |lambda|_max / |lambda|_min,nonzero
(1.9e4 at ny = 128 with tanh beta 1.5).  Two backward-stable FP64 solves of
the same system (the GPU's eigen-basis product and the oracle's per-mode LU)
may differ by kappa * eps relative, and the pressure is the running sum of
one increment per projection, so pressure (mean removed) and its compact
gradient are held to 4 * kappa * eps * (number of projections); kappa is
computed here from the oracle's own operators and printed.  Velocities are
held to 1e-12 (FP64) and 1e-6 (mixed) without such a factor.  The mixed
form's pressure has no accuracy target of its own: it is printed and only
checked to stay below 1e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)


from tolerances import EPS32, EPS64, ly_kappa, pressure_tol, rel_mean_removed, ulp_perturbed  # noqa: F401


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _grad_rel(of, pg, pr):
    """Relative difference of the compact pressure gradients (D1_C2P along x,
    y, z with the flow's boundary kinds, from the oracle's operators)."""
    dom = of.dom
    nx, ny, nz = dom.nc
    worst = 0.0
    for d, bc, shp in ((0, [1, 1], (nx, ny, nz)), (1, [5, 5], (nx, ny + 1, nz)), (2, [1, 1], (nx, ny, nz))):
        fbc = None
        if d == 1:
            fbc = np.zeros((nx, 4, nz), order="F")
        ga = dom.op3d(d, 4, 4, bc, pg, shp, fbc)
        gb = dom.op3d(d, 4, 4, bc, pr, shp, fbc)
        worst = max(worst, rel(ga, gb))
    return worst


def _floor_tols(of, n_proj, ref, pert):
    """pressure and pressure-gradient tolerances: the kappa bound, or ten
    times the oracle's own response to a one-ulp perturbation of the initial
    velocity (pert: its final state), whichever is larger"""
    kb = pressure_tol(of.dom, n_proj)
    sp = rel_mean_removed(pert[3], ref[3])
    sg = _grad_rel(of, pert[3], ref[3])
    print(f"pressure tolerance: kappa bound {kb:.1e}, oracle one-ulp response pr {sp:.1e} grad_p {sg:.1e}",
          flush=True)
    return max(kb, 10.0 * sp), max(kb, 10.0 * sg)


def _compare(tag, got, ref, of, vtol, ptol, gtol):
    errs = {}
    for name, g, r in zip(("qx", "qy", "qz"), got[:3], ref[:3]):
        errs[name] = rel(g, r)
    pg, pr = got[3] - got[3].mean(), ref[3] - ref[3].mean()
    errs["pr"] = rel(pg, pr)
    errs["grad_p"] = _grad_rel(of, got[3], ref[3])
    print(tag, " ".join(f"{k} {v:.2e}" for k, v in errs.items()),
          f"(tolerances: velocity {vtol:.1e}, pressure {ptol:.1e}, grad p {gtol:.1e})", flush=True)
    bad = [k for k in ("qx", "qy", "qz") if not errs[k] < vtol]
    bad += [k for k, t in (("pr", ptol), ("grad_p", gtol)) if not errs[k] < t]
    return errs, bad


def test_fp64_exact_full_size_vs_oracle(po):
    """256x128x256 channel, FP64 with the exact projection: initial step and
    two RK3 steps against the oracle at the benchmark's own size."""
    from chapsim2_amd.domain import PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (256, 128, 256)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=20261202)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **CFG)
    try:
        dm.set_state(*fields)
        dm.update_re(0)
        dm.substep(0)
        dm.run(0, 2)
        got = dm.get_state()
        div = dm.mass_check()
    finally:
        dm.close()
    finals = []
    for st in (fields, ulp_perturbed(fields)):
        of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
        of.set_state(*st)
        of.update_re(0)
        of.substep(0)
        for it in (1, 2):
            of.step(it)
        finals.append(of.get_state())
    pt, gt = _floor_tols(of, 7, finals[0], finals[1])
    print(f"kappa(L_y) {ly_kappa(of.dom):.3e}", flush=True)
    _, bad = _compare("256x128x256 fp64c 2 steps:", got, finals[0], of, 1e-12, pt, gt)
    print("max|div u|", div, flush=True)
    assert not bad, bad
    assert np.all(div < 1e-12)


@pytest.mark.parametrize("mixed", [False, True])
def test_channel_200_steps_vs_oracle(po, mixed):
    """The channel workload (CP6, tanh 1.5, Re_b 2800 constant flux, dt 5e-4,
    10 % noise) for 200 RK3 steps on a 32x24x16 grid: FP64 within 1e-12 of
    the oracle, the mixed form within 1e-6 per field; max|div u| at
    round-off throughout."""
    from chapsim2_amd.domain import PREC_FP64, PREC_MIXED, PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (32, 24, 16)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=2026)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, precision=PREC_MIXED if mixed else PREC_FP64,
                **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    op_ = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    try:
        dm.set_state(*fields)
        of.set_state(*fields)
        op_.set_state(*ulp_perturbed(fields))
        for x in (dm, of, op_):
            x.update_re(0)
            x.substep(0)
        for k in range(4):
            dm.run(50 * k, 50)
            for it in range(50 * k + 1, 50 * k + 51):
                of.step(it)
                op_.step(it)
            tag = f"{'mixed' if mixed else 'fp64c'} step {50 * (k + 1)}:"
            if mixed:
                _, bad = _compare(tag, dm.get_state(), of.get_state(), of, 1e-6, 1e-3, 1e-3)
            else:
                ref = of.get_state()
                pt, gt = _floor_tols(of, 1 + 3 * 50 * (k + 1), ref, op_.get_state())
                _, bad = _compare(tag, dm.get_state(), ref, of, 1e-12, pt, gt)
            assert not bad, (tag, bad)
            assert np.all(dm.mass_check() < 1e-12)
    finally:
        dm.close()


def test_tgv_64_100_steps_vs_oracle(po):
    """Taylor-Green vortex, 64^3, Re 1600, dt 1e-3, 3-D FFT Poisson: the
    initial step and all 100 RK3 steps of the benchmark case against the
    oracle."""
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain
    from chapsim2_amd.initial import tgv_fields

    nc = (64, 64, 64)
    L = 2 * np.pi
    cfg = dict(periodic_y=True, istret=0, dt=1e-3, ren=1600.0, idriven=0, poisson=FFT_2DECOMP_3DFFT)
    fields = tgv_fields(nc, L, -np.pi, np.pi, L)
    dm = Domain(nc, L, -np.pi, np.pi, L, **cfg)
    of = po.Flow(list(nc), L, -np.pi, np.pi, L, istret=0, wall_y=False, iacc=4, poisson=3, dt=1e-3, ren=1600.0,
                 idriven=0)
    try:
        dm.set_state(*fields)
        of.set_state(*fields)
        dm.update_re(0)
        of.update_re(0)
        dm.substep(0)
        of.substep(0)
        for k in range(4):
            dm.run(25 * k, 25)
            for it in range(25 * k + 1, 25 * k + 26):
                of.step(it)
            g, r = dm.get_state(), of.get_state()
            # qz starts at zero and grows by the nonlinearity: scaled by max |qx|
            errs = [rel(g[c], r[c]) for c in (0, 1)] + [float(np.abs(g[2] - r[2]).max() / np.abs(r[0]).max())]
            errs.append(rel(g[3] - g[3].mean(), r[3] - r[3].mean()))
            print(f"TGV 64^3 step {25 * (k + 1)}: qx {errs[0]:.2e} qy {errs[1]:.2e} qz {errs[2]:.2e} "
                  f"(max |qz| {np.abs(r[2]).max():.2e}) pr {errs[3]:.2e}", flush=True)
            assert max(errs) < 1e-12, errs
    finally:
        dm.close()


_BATCH = {}


def _batch_states(prec):
    """Final states of the 16 batched channels of the benchmark draw after one
    eager and two graph-captured RK3 steps (computed once per precision)."""
    if prec in _BATCH:
        return _BATCH[prec]
    import bench
    from chapsim2_amd.domain import Batch
    from chapsim2_amd.initial import channel_fields

    specs = bench.batch_specs(16)
    for s in specs:
        s["precision"] = prec
    b = Batch(specs)
    try:
        fields = []
        for i, (d, s) in enumerate(zip(b.domains, specs)):
            f = channel_fields(s["nc"], s["lxx"], -1.0, 1.0, s["lzz"], 2, s["rstret"], noise=0.1, seed=100 + i)
            fields.append(f)
            d.set_state(*f)
            d.update_re(0)
            d.substep(0)
        b.run(0, 3, graph=True)
        states = [d.get_state() for d in b.domains]
        divs = [float(d.mass_check().max()) for d in b.domains]
    finally:
        b.close()
    _BATCH[prec] = (specs, fields, states, divs)
    return _BATCH[prec]


@pytest.mark.parametrize("i", range(16))
def test_batched_channels_vs_oracle(po, i):
    """Instance i of the 16-channel batched benchmark (grids 64^3 ..
    192x128x192, x/2 and z lengths such as 36, 72, 90, 100, 144, 180 with
    3^2 and 5^2 factors, Re_b 2800 / 6900, tanh beta in [1, 2]) after one
    eager and two graph-captured RK3 steps: the FP64 batch within 1e-12 of
    the oracle, the mixed batch (the benchmark's form) within 1e-6."""
    from chapsim2_amd.domain import PREC_FP64, PREC_MIXED

    specs, fields, s64, d64 = _batch_states(PREC_FP64)
    _, _, s32, d32 = _batch_states(PREC_MIXED)
    s = specs[i]
    finals = []
    for st in (fields[i], ulp_perturbed(fields[i])):
        of = po.Flow(list(s["nc"]), s["lxx"], -1.0, 1.0, s["lzz"], wall_y=True, iacc=4, poisson=2, projection=1,
                     istret=2, rstret=s["rstret"], dt=s["dt"], ren=s["ren"], idriven=1)
        of.set_state(*st)
        of.update_re(0)
        of.substep(0)
        for it in (1, 2, 3):
            of.step(it)
        finals.append(of.get_state())
    ref = finals[0]
    pt, gt = _floor_tols(of, 10, ref, finals[1])
    print(f"kappa(L_y) {ly_kappa(of.dom):.3e}", flush=True)
    _, bad64 = _compare(f"batch[{i}] {s['nc']} fp64:", s64[i], ref, of, 1e-12, pt, gt)
    _, bad32 = _compare(f"batch[{i}] {s['nc']} mixed:", s32[i], ref, of, 1e-6, 1e-3, 1e-3)
    assert not bad64 and not bad32, (bad64, bad32)
    assert d64[i] < 1e-12 and d32[i] < 1e-12


def test_exact_projection_multi_tile_ymix(po):
    """(256, 24, 256): 1032 y-mixing tiles of 64 spectrum columns (more
    workgroups than the CUs hold at once, and ny / 2 = 12 rows of the
    64-row blocks in use); the initial step and one RK3 step against the
    oracle's per-mode dense solve."""
    from chapsim2_amd.domain import PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (256, 24, 256)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=77)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    op_ = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    try:
        for x, st in ((dm, fields), (of, fields), (op_, ulp_perturbed(fields))):
            x.set_state(*st)
            x.update_re(0)
            x.substep(0)
        dm.run(0, 1)
        of.step(1)
        op_.step(1)
        ref = of.get_state()
        pt, gt = _floor_tols(of, 4, ref, op_.get_state())
        _, bad = _compare("256x24x256 fp64c multi-tile:", dm.get_state(), ref, of, 1e-12, pt, gt)
        assert not bad, bad
        assert np.all(dm.mass_check() < 1e-12)
    finally:
        dm.close()


@pytest.mark.parametrize("op", range(6))
def test_fp32_x_operators_512_rows_vs_oracle(po, op):
    """The FP32 x-line kernels at 512 rows (the 512x256x512 operator
    benchmark's x lines; here 512 x 16 x 32, launch shapes autotuned as in
    the benchmark) against the FP64 oracle on band-limited lines.  The
    operator is circulant, so a relative perturbation of size eps of input,
    coefficients and arithmetic moves an output by at most
    eps * ||M||_inf * max|f|, with ||M||_inf the sum of the magnitudes of
    its impulse response (1.67 for the midpoint interpolations, 106-112
    for the first derivatives at this spacing): the error is held to
    4 * EPS32 * ||M||_inf * max|f| / max|ref|, and above FP64 level."""
    import ctypes as C

    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import Domain

    nc = (512, 16, 32)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, **CFG)
    odm = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    try:
        rng = np.random.default_rng(op)
        x = np.arange(nc[0]) * 2 * np.pi / nc[0]
        ph = rng.uniform(0.0, 2 * np.pi, (1, nc[1], nc[2]))
        fi = np.asfortranarray(sum(np.cos(m * x[:, None, None] + m * ph) / m for m in (1, 5, 23, 101, 170)))
        ref = odm.op3d(0, op, 4, [1, 1], fi, list(nc))
        ti = torch.from_numpy(np.ascontiguousarray(fi.T)).float().cuda()
        to = torch.zeros(nc[::-1], dtype=torch.float32, device="cuda")
        _lib.check(dm.lib.chs_op3d_f32(dm.h, 0, op, 4, _lib.ivec([1, 1]), C.c_void_p(ti.data_ptr()),
                                       _lib.ivec(nc), C.c_void_p(to.data_ptr()), _lib.ivec(nc), None))
        got = to.double().cpu().numpy().T
        err = rel(got, ref)
        imp = np.zeros(nc[0])
        imp[nc[0] // 2] = 1.0
        mnorm = float(np.abs(odm.op1d(0, op, 4, [1, 1], imp, nc[0])).sum())
        tol = 4 * EPS32 * mnorm * np.abs(fi).max() / np.abs(ref).max()
        print(f"x op {op} at 512 rows, FP32 vs FP64 oracle: {err:.2e} (||M||_inf {mnorm:.1f}, "
              f"tolerance {tol:.2e})", flush=True)
        assert 1e-9 < err < tol
    finally:
        dm.close()


@pytest.mark.parametrize("ny", [128, 150])
@pytest.mark.parametrize("op", range(6))
def test_fp32_y_operators_long_chunks_vs_oracle(po, ny, op):
    """FP32 y lines of 129-160 output rows run on 32-row chunks (5 instead of
    9 chunks of 16); every operator against the FP64 oracle, walls with the
    operator's boundary kinds (Neumann for cell-centred inputs, Dirichlet for
    node inputs, zero wall values).  A relative perturbation eps of input,
    coefficients and arithmetic moves an output by at most
    eps * ||M||_inf * max|f| with ||M||_inf the largest absolute row sum of
    the operator's matrix (assembled from the oracle's impulse responses);
    the error is held to 4 * EPS32 * ||M||_inf * max|f| / max|ref|."""
    import ctypes as C

    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import Domain

    nc = (16, ny, 8)
    cin = op in (0, 2, 4)  # MIDP_C2P, D1_C2C, D1_C2P read cell-centred values
    cout = op in (1, 2, 5)
    bc = [5, 5] if cin else [4, 4]
    n_in = ny if cin else ny + 1
    n_out = ny if cout else ny + 1
    ei, eo = [nc[0], n_in, nc[2]], [nc[0], n_out, nc[2]]
    dm = Domain(nc, LX, -1.0, 1.0, LZ, **CFG)
    odm = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    try:
        rng = np.random.default_rng(100 + op + ny)
        s = np.linspace(-1.0, 1.0, n_in)
        ph = rng.uniform(0.0, 2 * np.pi, (nc[0], 1, nc[2]))
        fi = np.asfortranarray(sum(np.cos(k * np.pi * s[None, :, None] + ph) / k for k in (1, 3, 7, 19)))
        fbc = np.zeros((nc[0], 4, nc[2]), order="F")
        ref = odm.op3d(1, op, 4, bc, fi, eo, fbc)
        ti = torch.from_numpy(np.ascontiguousarray(fi.T)).float().cuda()
        to = torch.zeros(eo[::-1], dtype=torch.float32, device="cuda")
        tb = torch.zeros(nc[2], 4, nc[0], dtype=torch.float32, device="cuda")
        _lib.check(dm.lib.chs_op3d_f32(dm.h, 1, op, 4, _lib.ivec(bc), C.c_void_p(ti.data_ptr()), _lib.ivec(ei),
                                       C.c_void_p(to.data_ptr()), _lib.ivec(eo), C.c_void_p(tb.data_ptr())))
        got = to.double().cpu().numpy().T
        err = rel(got, ref)
        M = np.stack([odm.op1d(1, op, 4, bc, np.eye(n_in)[j], n_out, np.zeros(4)) for j in range(n_in)], axis=1)
        mnorm = float(np.abs(M).sum(axis=1).max())
        tol = 4 * EPS32 * mnorm * np.abs(fi).max() / np.abs(ref).max()
        print(f"y op {op}, {n_out} rows, FP32 vs FP64 oracle: {err:.2e} (||M||_inf {mnorm:.1f}, "
              f"tolerance {tol:.2e})", flush=True)
        assert 1e-9 < err < tol
    finally:
        dm.close()
