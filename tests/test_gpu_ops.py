"""GPU parity of the 18 compact operators (chs_op3d) against the oracle's
restatement of Get_*_3D (src/basics_operations2.f90:3952-4790)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NC = (32, 24, 16)
LX, LZ = 4 * np.pi, 4 * np.pi / 3
# FP64 tolerance, relative to the largest output magnitude
RTOL = 1e-12

INP = {0: "c", 1: "p", 2: "c", 3: "p", 4: "c", 5: "p"}
OUT = {0: "p", 1: "c", 2: "c", 3: "p", 4: "p", 5: "c"}


@pytest.fixture(scope="module")
def chan():
    from chapsim2_amd.domain import Domain

    dm = Domain(NC, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    yield dm
    dm.close()


def _n(ncd, per, loc):
    return ncd if (per or loc == "c") else ncd + 1


def _cases():
    for d in (0, 1, 2):
        for op in range(6):
            for acc in (1, 2, 3, 4):
                if d != 1:
                    yield d, op, acc, (1, 1), False
                else:
                    for b1 in (0, 2, 3, 4, 5, 6):
                        for b2 in (0, 2, 3, 4, 5, 6):
                            for fb in (False, True):
                                if (b1 == 0 or b2 == 0) and not fb:
                                    continue
                                yield d, op, acc, (b1, b2), fb


def test_ops_channel(chan, po):
    from chapsim2_amd import operations as ops

    odm = po.Dom(list(NC), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    rng = np.random.default_rng(7)
    nchecked, worst = 0, 0.0
    for d, op, acc, bc, fb in _cases():
        per = bc[0] == 1
        ext_in = list(NC)
        ext_out = list(NC)
        ext_in[d] = _n(NC[d], per, INP[op])
        ext_out[d] = _n(NC[d], per, OUT[op])
        fi = np.asfortranarray(rng.standard_normal(ext_in))
        fshape = list(ext_in)
        fshape[d] = 4
        fbc = np.asfortranarray(rng.standard_normal(fshape)) if fb else None
        try:
            ref = odm.op3d(d, op, acc, list(bc), fi, ext_out, fbc)
        except ValueError:
            continue
        if not np.all(np.isfinite(ref)) or np.abs(ref).max() > 1e15:
            continue  # combinations whose reference ghost value is undefined
        out = np.zeros(ext_out, order="F")
        ops.op3d(chan, d, op, fi, out, acc, list(bc), fbc)
        err = np.abs(out - ref).max() / max(1.0, np.abs(ref).max())
        worst = max(worst, err)
        assert err < RTOL, (d, op, acc, bc, fb, err)
        nchecked += 1
    assert nchecked > 500
    print("checked", nchecked, "worst", worst)


def test_named_operators_match_generic(chan, po):
    """Get_y_1der_C2P_3D et al. are the generic op with (dir, kind) bound."""
    from chapsim2_amd import operations as ops

    rng = np.random.default_rng(3)
    nx, ny, nz = NC
    fi = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    fbc = np.asfortranarray(np.zeros((nx, 4, nz)))
    a = np.zeros((nx, ny + 1, nz), order="F")
    ops.Get_y_1der_C2P_3D(fi, a, chan, 4, [4, 4], fbc)
    odm = po.Dom(list(NC), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    ref = odm.op3d(1, 4, 4, [4, 4], fi, [nx, ny + 1, nz], fbc)
    assert np.abs(a - ref).max() / np.abs(ref).max() < RTOL


def test_ops_linearity_on_gpu(chan):
    """The y-line C2P derivative with Dirichlet walls is linear in its input
    at full size: op(a f + b g) == a op(f) + b op(g) (wall values zero)."""
    from chapsim2_amd import operations as ops

    rng = np.random.default_rng(11)
    nx, ny, nz = NC
    f = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    g = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    outs = []
    for arr in (f, g, 2.0 * f - 3.0 * g):
        o = np.zeros((nx, ny + 1, nz), order="F")
        ops.op3d(chan, 1, 4, np.asfortranarray(arr), o, 4, [4, 4], np.zeros((nx, 4, nz), order="F"))
        outs.append(o)
    lhs = outs[2]
    rhs = 2.0 * outs[0] - 3.0 * outs[1]
    assert np.abs(lhs - rhs).max() / np.abs(rhs).max() < 1e-12


NC_LONG = (250, 130, 256)


def test_ops_long_lines(po):
    """Line lengths 130-257 (9 to 16 chunks per line in the partitioned
    kernels, uneven chunk lengths for 130/131/250) against the oracle."""
    from chapsim2_amd import operations as ops
    from chapsim2_amd.domain import Domain

    dm = Domain(NC_LONG, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    odm = po.Dom(list(NC_LONG), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    rng = np.random.default_rng(17)
    nchecked = 0
    try:
        for d in (0, 1, 2):
            for op in range(6):
                for acc in (1, 4):
                    bcs = [((1, 1), False)] if d != 1 else [((4, 4), True), ((5, 5), True), ((2, 3), False)]
                    for bc, fb in bcs:
                        per = bc[0] == 1
                        ext_in = list(NC_LONG)
                        ext_out = list(NC_LONG)
                        ext_in[d] = _n(NC_LONG[d], per, INP[op])
                        ext_out[d] = _n(NC_LONG[d], per, OUT[op])
                        fi = np.asfortranarray(rng.standard_normal(ext_in))
                        fshape = list(ext_in)
                        fshape[d] = 4
                        fbc = np.asfortranarray(rng.standard_normal(fshape)) if fb else None
                        try:
                            ref = odm.op3d(d, op, acc, list(bc), fi, ext_out, fbc)
                        except ValueError:
                            continue
                        if not np.all(np.isfinite(ref)) or np.abs(ref).max() > 1e15:
                            continue
                        out = np.zeros(ext_out, order="F")
                        ops.op3d(dm, d, op, fi, out, acc, list(bc), fbc)
                        err = np.abs(out - ref).max() / max(1.0, np.abs(ref).max())
                        assert err < RTOL, (d, op, acc, bc, fb, err)
                        nchecked += 1
    finally:
        dm.close()
    assert nchecked > 40


@pytest.mark.parametrize("nc,d", [((512, 8, 24), 0), ((300, 8, 12), 0), ((16, 400, 8), 1), ((16, 257, 8), 1),
                                  ((16, 511, 8), 1), ((8, 12, 512), 2), ((16, 8, 450), 2)])
def test_ops_long_chunk_lines(po, nc, d):
    """Lines of 257-512 rows take the partitioned kernels with chunks of up
    to 32 rows (x lines with the lanes one line apart, y and z lines with
    the lanes along x; y lines of 512 p-located rows exceed 16 chunks and
    take the single-thread-per-line kernel); every operator and accuracy,
    periodic x and z, wall-bounded y with Dirichlet, Neumann and interior
    boundary kinds, against the oracle."""
    from chapsim2_amd import operations as ops
    from chapsim2_amd.domain import Domain

    dm = Domain(nc, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    odm = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    rng = np.random.default_rng(23 + d)
    nchecked = 0
    try:
        for op in range(6):
            for acc in (1, 2, 3, 4):
                bcs = [((1, 1), False)] if d != 1 else [((4, 4), True), ((5, 5), True), ((2, 3), False),
                                                         ((0, 6), True)]
                for bc, fb in bcs:
                    per = bc[0] == 1
                    ext_in, ext_out = list(nc), list(nc)
                    ext_in[d] = _n(nc[d], per, INP[op])
                    ext_out[d] = _n(nc[d], per, OUT[op])
                    fi = np.asfortranarray(rng.standard_normal(ext_in))
                    fshape = list(ext_in)
                    fshape[d] = 4
                    fbc = np.asfortranarray(rng.standard_normal(fshape)) if fb else None
                    try:
                        ref = odm.op3d(d, op, acc, list(bc), fi, ext_out, fbc)
                    except ValueError:
                        continue
                    if not np.all(np.isfinite(ref)) or np.abs(ref).max() > 1e15:
                        continue
                    out = np.zeros(ext_out, order="F")
                    ops.op3d(dm, d, op, fi, out, acc, list(bc), fbc)
                    err = np.abs(out - ref).max() / max(1.0, np.abs(ref).max())
                    assert err < RTOL, (nc, d, op, acc, bc, fb, err)
                    nchecked += 1
    finally:
        dm.close()
    assert nchecked >= 24


def test_ops_fp32_vs_oracle(chan, po):
    """FP32 operators (chs_op3d_f32, the kernels of the mixed-precision
    solver) against the FP64 oracle: every operator along every direction
    with the channel boundary kinds, error relative to the largest output
    within 2e-6 (FP32 round-off, 6e-8, amplified by the compact operators'
    stencil sums and line solves)."""
    import ctypes as C

    import torch

    from chapsim2_amd import _lib

    odm = po.Dom(list(NC), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
    rng = np.random.default_rng(29)
    worst = 0.0
    for d in (0, 1, 2):
        for op in range(6):
            bc = [1, 1] if d != 1 else ([5, 5] if INP[op] == "c" else [4, 4])
            ext_in, ext_out = list(NC), list(NC)
            ext_in[d] = _n(NC[d], d != 1, INP[op])
            ext_out[d] = _n(NC[d], d != 1, OUT[op])
            fi = np.asfortranarray(rng.standard_normal(ext_in))
            fbc = None
            if d == 1:
                fshape = list(ext_in)
                fshape[1] = 4
                fbc = np.asfortranarray(rng.standard_normal(fshape))
            ref = odm.op3d(d, op, 4, bc, fi, ext_out, fbc)
            ti = torch.from_numpy(np.ascontiguousarray(fi.T)).float().cuda()
            to = torch.zeros(ext_out[::-1], dtype=torch.float32, device="cuda")
            tb = torch.from_numpy(np.ascontiguousarray(fbc.T)).float().cuda() if fbc is not None else None
            _lib.check(chan.lib.chs_op3d_f32(chan.h, d, op, 4, _lib.ivec(bc), C.c_void_p(ti.data_ptr()),
                                             _lib.ivec(ext_in), C.c_void_p(to.data_ptr()), _lib.ivec(ext_out),
                                             C.c_void_p(tb.data_ptr()) if tb is not None else None))
            got = to.double().cpu().numpy().T
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"f32 op dir {d} op {op}: {err:.2e}")
            assert err < 2e-6, (d, op, err)
            worst = max(worst, err)
    assert worst > 1e-9  # the arithmetic really was FP32


def test_poisson_fp32_vs_oracle(chan, po):
    """FP32 Poisson solve (chs_poisson_f32) against the oracle's FP64
    FFTPACK-form solve of the same zero-mean right-hand side."""
    import ctypes as C

    import torch

    from chapsim2_amd import _lib

    of = po.Flow(list(NC), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, istret=2, rstret=0.075, dt=5e-4,
                 ren=2800.0, idriven=1)
    rng = np.random.default_rng(31)
    rhs = rng.standard_normal(NC)
    rhs -= rhs.mean()
    rhs = np.asfortranarray(rhs)
    ref = of.poisson(rhs)
    t = torch.from_numpy(np.ascontiguousarray(rhs.T)).float().cuda()
    _lib.check(chan.lib.chs_poisson_f32(chan.h, C.c_void_p(t.data_ptr())))
    got = t.double().cpu().numpy().T
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"f32 poisson: {err:.2e}")
    assert 1e-9 < err < 1e-5, err


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_x_launch_shapes_vs_oracle(po, variant):
    """Every x-line launch shape, forced through chs_tune_force (0: 64 whole
    lines in LDS, 1: 32 whole lines in LDS with two chunks per wave, 2: skewed
    rows, 3: lane-per-line, long lines only, 4: 16 whole lines in LDS with
    four chunks per wave, 16-row chunks only, 5: persistent workgroups over
    32-line tiles with the next tile's input prefetched, one or two input
    arrays read 16 bytes at a time; a shape that does not apply falls back to
    the next that does): all operators on lines of 250 rows (16-row chunks),
    300 and 512 rows (32-row chunks), CP6 and the explicit CD4, then the
    momentum right-hand side (merged operators reading two to four arrays
    along x), against the oracle.  Shape 5 also runs 256-row lines on 9 000
    lines (282 tiles: more tiles than workgroups, a last tile of 8 lines)."""
    from chapsim2_amd import _lib
    from chapsim2_amd import operations as ops
    from chapsim2_amd.domain import Domain

    lib = _lib.load()
    _lib.check(lib.chs_tune_force(b"line_x/", variant))
    try:
        grids = ((250, 8, 12), (300, 8, 12), (512, 8, 8))
        if variant == 5:
            grids = grids + ((256, 90, 100),)
        for nc in grids:
            dm = Domain(nc, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
            odm = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
            rng = np.random.default_rng(41 + variant)
            try:
                for op in range(6):
                    for acc in (2, 4):
                        fi = np.asfortranarray(rng.standard_normal(nc))
                        ref = odm.op3d(0, op, acc, [1, 1], fi, list(nc), None)
                        out = np.zeros(nc, order="F")
                        ops.op3d(dm, 0, op, fi, out, acc, [1, 1], None)
                        err = np.abs(out - ref).max() / np.abs(ref).max()
                        assert err < RTOL, (variant, nc, op, acc, err)
                if variant == 5 and nc[0] == 256:
                    # the FP32 kernels of the same shape (four 16-byte groups per thread)
                    import ctypes as C

                    import torch

                    for op in range(6):
                        fi = np.asfortranarray(rng.standard_normal(nc))
                        ref = odm.op3d(0, op, 4, [1, 1], fi, list(nc), None)
                        ti = torch.from_numpy(np.ascontiguousarray(fi.T)).float().cuda()
                        to = torch.zeros(nc[::-1], dtype=torch.float32, device="cuda")
                        _lib.check(dm.lib.chs_op3d_f32(dm.h, 0, op, 4, _lib.ivec([1, 1]), C.c_void_p(ti.data_ptr()),
                                                       _lib.ivec(list(nc)), C.c_void_p(to.data_ptr()),
                                                       _lib.ivec(list(nc)), None))
                        got = to.double().cpu().numpy().T
                        err = np.abs(got - ref).max() / np.abs(ref).max()
                        assert 1e-9 < err < 2e-6, (variant, "f32", op, err)
            finally:
                dm.close()
        from chapsim2_amd.initial import channel_fields

        cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
        dm = Domain(NC, LX, -1.0, 1.0, LZ, **cfg)
        of = po.Flow(list(NC), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **cfg)
        try:
            fields = channel_fields(NC, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=777)
            dm.set_state(*fields)
            of.set_state(*fields)
            for isub in (0, 1):
                for g, r, name in zip(dm.momentum_rhs(isub), of.momentum_rhs(isub), ("mx", "my", "mz")):
                    err = np.abs(g - r).max() / np.abs(r).max()
                    assert err < RTOL, (variant, isub, name, err)
        finally:
            dm.close()
    finally:
        _lib.check(lib.chs_tune_force(b"line_x/", -1))


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_yz_pair_shapes_vs_oracle(po, variant):
    """The FP32 y/z operator pairs (interpolation and derivative of one
    input), forced through chs_tune_force (0: one pass over 64 lines per
    workgroup, 1: the two operators one by one, 2: one pass over 32 lines
    per workgroup; FP64 pairs always take the 64-line pass): on channels
    with 24- and 129-row y lines and 16- and 256-row z lines, the FP64
    momentum right-hand side, whose first level runs every pair, against
    the oracle (1e-12), and the mixed form's (16- and 32-row chunks) in
    each forced shape against the 64-line pass (1e-6)."""
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import PREC_MIXED, Domain
    from chapsim2_amd.initial import channel_fields

    lib = _lib.load()
    cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    try:
        for nc in ((40, 24, 16), (24, 129, 256)):
            print(f"pairs variant {variant} nc={nc}", flush=True)
            fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=17)
            of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, **cfg)
            of.set_state(*fields)
            ref = of.momentum_rhs(1)
            _lib.check(lib.chs_tune_force(b"line_yz2/", variant))
            dm = Domain(nc, LX, -1.0, 1.0, LZ, **cfg)
            try:
                dm.set_state(*fields)
                got = dm.momentum_rhs(1)
            finally:
                dm.close()
            for name, g, r in zip(("mx", "my", "mz"), got, ref):
                err = np.abs(g - r).max() / np.abs(r).max()
                assert err < 1e-12, (variant, nc, name, err)
            rhs = []
            for v in (0, variant):
                _lib.check(lib.chs_tune_force(b"line_yz2/", v))
                dm = Domain(nc, LX, -1.0, 1.0, LZ, precision=PREC_MIXED, **cfg)
                try:
                    dm.set_state(*fields)
                    rhs.append(dm.momentum_rhs(1))
                finally:
                    dm.close()
            for name, a, b in zip(("mx", "my", "mz"), *rhs):
                err = np.abs(a - b).max() / np.abs(b).max()
                assert err < 1e-6, (variant, nc, name, err)
    finally:
        _lib.check(lib.chs_tune_force(b"line_yz2/", -1))


@pytest.mark.parametrize("variant", [0, 1])
def test_yz_launch_shapes_vs_oracle(po, variant):
    """Both y/z launch shapes, forced through chs_tune_force (0: 64 lines per
    workgroup, one chunk per wave; 1: 32 lines per workgroup, two chunks per
    wave, single-input operators and every FP32 operator), on y lines of
    24/25 and 129/130 rows and z lines of 16 and 256 rows, every operator with
    CP6 and the explicit CD4, wall-bounded y with Dirichlet, Neumann and
    interpolation boundary kinds, against the oracle; then the FP32 merged
    operators of the mixed momentum right-hand side against the 64-line
    shape."""
    from chapsim2_amd import _lib
    from chapsim2_amd import operations as ops
    from chapsim2_amd.domain import Domain

    lib = _lib.load()
    _lib.check(lib.chs_tune_force(b"line_yz/", variant))
    try:
        for nc in ((40, 24, 16), (24, 129, 256)):
            dm = Domain(nc, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
            odm = po.Dom(list(nc), [1, 0, 1], LX, -1.0, 1.0, LZ, 2, 0.075)
            rng = np.random.default_rng(61 + variant)
            try:
                for d in (1, 2):
                    for op in range(6):
                        for acc in (2, 4):
                            bcs = [((1, 1), False)] if d != 1 else [((4, 4), True), ((5, 5), True), ((2, 3), False)]
                            for bc, fb in bcs:
                                per = bc[0] == 1
                                ext_in, ext_out = list(nc), list(nc)
                                ext_in[d] = _n(nc[d], per, INP[op])
                                ext_out[d] = _n(nc[d], per, OUT[op])
                                fi = np.asfortranarray(rng.standard_normal(ext_in))
                                fshape = list(ext_in)
                                fshape[d] = 4
                                fbc = np.asfortranarray(rng.standard_normal(fshape)) if fb else None
                                try:
                                    ref = odm.op3d(d, op, acc, list(bc), fi, ext_out, fbc)
                                except ValueError:
                                    continue
                                if not np.all(np.isfinite(ref)) or np.abs(ref).max() > 1e15:
                                    continue
                                out = np.zeros(ext_out, order="F")
                                ops.op3d(dm, d, op, fi, out, acc, list(bc), fbc)
                                err = np.abs(out - ref).max() / max(1.0, np.abs(ref).max())
                                assert err < RTOL, (variant, nc, d, op, acc, bc, fb, err)
            finally:
                dm.close()
        # FP32 merged operators (two to four input arrays: the mixed form's
        # momentum right-hand side) in this shape against the 64-line shape
        from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT
        from chapsim2_amd.initial import channel_fields

        nc = (32, 40, 48)
        fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=5)
        rhs = []
        for v in (0, variant):
            _lib.check(lib.chs_tune_force(b"line_yz/", v))
            dm = Domain(nc, LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1,
                        projection=PROJ_EXACT, precision=PREC_MIXED)
            try:
                dm.set_state(*fields)
                rhs.append(dm.momentum_rhs(1))
            finally:
                dm.close()
        for name, a, b in zip(("mx", "my", "mz"), *rhs):
            err = np.abs(a - b).max() / np.abs(b).max()
            assert err < 1e-6, (variant, name, err)
    finally:
        _lib.check(lib.chs_tune_force(b"line_yz/", -1))
