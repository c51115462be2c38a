"""Pins the oracle's FFTPACK-convention transforms
(lib/fishpack4.1/fftpack.f:129-240), its FISHPACK Poisson solve
(src/poisson_fishpack.f90:20-57, 213-490) and its exact projection."""
import numpy as np
import pytest


def _fftpack_forward(r):
    """RFFTF definition: r[0] = sum, then (Re, Im) pairs of sum r e^{-i...}
    for wavenumbers 1..n/2 (the last one real for even n)."""
    n = len(r)
    F = np.fft.rfft(r)
    out = np.empty(n)
    out[0] = F[0].real
    for k in range(1, (n - 1) // 2 + 1):
        out[2 * k - 1] = F[k].real
        out[2 * k] = F[k].imag
    if n % 2 == 0:
        out[n - 1] = F[n // 2].real
    return out


@pytest.mark.parametrize("n", [8, 12, 30, 64, 96])
def test_rfft_against_numpy(po, n):
    rng = np.random.default_rng(n)
    r = rng.standard_normal(n)
    f = po.rfftf(r)
    assert np.abs(f - _fftpack_forward(r)).max() < 1e-12 * n
    assert np.abs(po.rfftb(f) - n * r).max() < 1e-12 * n


def _fishpack_operator(f):
    """Dense 2nd-order operator of fishpack_fft_init (x, z periodic; y walls
    with Neumann closure), built from the mesh alone."""
    dm = f.dom
    nx, ny, nz = dm.nc
    hx, hz = dm.h[0], dm.h[2]
    def per(n, h):
        D = -2.0 * np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)
        D[0, -1] = D[-1, 0] = 1.0
        return D / h**2
    ymc, ymp = dm.ymc, dm.ymp
    hy = dm.h[1]
    Y = np.zeros((ny, ny))
    for j in range(ny):
        a = ymc[j] * ymp[j] / hy**2
        c = ymc[j] * ymp[j + 1] / hy**2
        if j > 0:
            Y[j, j - 1] += a
            Y[j, j] -= a
        if j < ny - 1:
            Y[j, j + 1] += c
            Y[j, j] -= c
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    # Fortran order (x fastest): index = i + nx*(j + ny*k)
    return (np.kron(Iz, np.kron(Iy, per(nx, hx))) + np.kron(Iz, np.kron(Y, Ix))
            + np.kron(per(nz, hz), np.kron(Iy, Ix)))


def test_fishpack_solves_2nd_order_operator(po):
    nc = [8, 10, 6]
    f = po.Flow(nc, 2.0, -1.0, 1.0, 1.5, istret=2, rstret=0.075, iacc=1, poisson=2)
    A = _fishpack_operator(f)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(A.shape[0])
    rhs = A @ x  # compatible right-hand side
    phi = f.poisson(np.asfortranarray(rhs.reshape(nc, order="F"))).reshape(-1, order="F")
    assert np.abs(A @ phi - rhs).max() < 1e-10 * np.abs(rhs).max()
    # gauge of TRID0's singular pivot: the solution differs from x by a constant
    d = phi - x
    assert np.ptp(d) < 1e-10 * np.abs(x).max()


@pytest.mark.parametrize("iacc", [1, 2, 3, 4])
def test_exact_projection_zero_divergence(po, iacc):
    """One exact projection leaves the discrete divergence at round-off, and
    a second one changes nothing."""
    from chapsim2_amd.initial import channel_fields

    nc = (24, 20, 12)
    lx, lz = 2 * np.pi, np.pi
    cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1, iacc=iacc)
    fields = channel_fields(nc, lx, -1.0, 1.0, lz, 2, 0.075, noise=0.1, seed=3)
    f = po.Flow(list(nc), lx, -1.0, 1.0, lz, wall_y=True, poisson=2, projection=1, **cfg)
    f.set_state(*fields)
    f.update_re(0)
    d0 = np.abs(f.divergence()).max()
    f.project(1)
    d1 = np.abs(f.divergence()).max()
    q1 = f.get_state()
    f.project(1)
    q2 = f.get_state()
    assert d0 > 1e-2 and d1 < 1e-13 * max(1.0, d0)
    for a, b in zip(q1[:3], q2[:3]):
        assert np.abs(a - b).max() < 1e-13


def test_exact_projection_equals_fishpack_for_cd2(po):
    """With 2nd-order central schemes the composed D*G is FISHPACK's
    operator, so the exact projection and the reference projection give the
    same velocities (an independent pin of both solves)."""
    from chapsim2_amd.initial import channel_fields

    nc = (24, 20, 12)
    lx, lz = 2 * np.pi, np.pi
    cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1, iacc=1)
    fields = channel_fields(nc, lx, -1.0, 1.0, lz, 2, 0.075, noise=0.1, seed=4)
    out = []
    for proj in (0, 1):
        f = po.Flow(list(nc), lx, -1.0, 1.0, lz, wall_y=True, poisson=2, projection=proj, **cfg)
        f.set_state(*fields)
        f.update_re(0)
        f.project(1)
        out.append(f.get_state())
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(a).max())
    # pressures agree up to the gauge constant
    dp = out[0][3] - out[1][3]
    assert np.ptp(dp) < 1e-9 * max(1.0, np.abs(out[0][3]).max())


def test_reference_projection_growth_with_cp6(po):
    """The reference scheme (2nd-order Poisson, CP6 D and G) does not remove
    the divergence: repeated projections contract it only slowly, and RK3
    steps amplify it.  The exact projection does not have this problem."""
    from chapsim2_amd.initial import channel_fields

    nc = (32, 24, 16)
    lx, lz = 4 * np.pi, 4 * np.pi / 3
    cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1, iacc=4)
    fields = channel_fields(nc, lx, -1.0, 1.0, lz, 2, 0.075, noise=0.1, seed=7)
    divs = {}
    for proj in (0, 1):
        f = po.Flow(list(nc), lx, -1.0, 1.0, lz, wall_y=True, poisson=2, projection=proj, **cfg)
        f.set_state(*fields)
        f.update_re(0)
        f.substep(0)
        for it in range(1, 4):
            f.step(it)
        divs[proj] = np.abs(f.divergence()).max()
    assert divs[0] > 1.0
    assert divs[1] < 1e-13


def test_poisson_000_single_mode(po):
    """Triply periodic path: a Fourier mode is divided by minus the squared
    modified wavenumbers of the staggered CP6 derivative (published
    coefficients alpha = 9/62, a = 63/62, b = 17/186 per unit spacing)."""
    nc = [16, 12, 10]
    L = 2 * np.pi
    f = po.Flow(nc, L, -np.pi, np.pi, L, istret=0, wall_y=False, iacc=4, poisson=3)
    h = [L / n for n in nc]
    xc = (np.arange(nc[0]) + 0.5) * h[0]
    yc = -np.pi + (np.arange(nc[1]) + 0.5) * h[1]
    zc = (np.arange(nc[2]) + 0.5) * h[2]
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    m = (3, 2, 4)
    rhs = np.cos(m[0] * X) * np.cos(m[1] * Y) * np.sin(m[2] * Z)

    def kp(mm, n):
        w = 2 * np.pi * mm / n
        al, a, b = 9 / 62, 63 / 62, 17 / 186
        return n / L * (2 * a * np.sin(w / 2) + 2 * b * np.sin(1.5 * w)) / (1 + 2 * al * np.cos(w))

    k2 = sum(kp(mm, n) ** 2 for mm, n in zip(m, nc))
    phi = f.poisson(np.asfortranarray(rhs))
    assert np.abs(phi + rhs / k2).max() < 1e-13 * np.abs(rhs / k2).max()
    # the mean is removed
    phi0 = f.poisson(np.asfortranarray(np.ones(nc)))
    assert np.abs(phi0).max() < 1e-14


@pytest.mark.parametrize("iacc,nc", [(4, (16, 12, 10)), (4, (12, 16, 15)), (1, (10, 8, 12)), (3, (8, 20, 9))])
def test_poisson_010_is_poisson_000_of_the_mirrored_channel(po, iacc, nc):
    """poisson_010 (uniform channel, y walls): the right-hand side mirrored
    across a wall (2 ny cells over 2 ly, periodic) and solved by the pinned
    triply periodic poisson_000 gives, on the first ny rows, the channel
    solution. The half-cell mirror is the cosine series of the reference's
    reordered transform, and both divide by the same modified wavenumbers,
    so a slip in the reorder, the y post-processing, the z half spectrum or
    the walled y wavenumbers breaks the equality (odd nz included)."""
    nx, ny, nz = nc
    lx, lz = 2.0, 1.5
    ch = po.Flow(list(nc), lx, -1.0, 1.0, lz, istret=0, wall_y=True, iacc=iacc, poisson=3)
    box = po.Flow([nx, 2 * ny, nz], lx, -1.0, 3.0, lz, istret=0, wall_y=False, iacc=iacc, poisson=3)
    rng = np.random.default_rng(sum(nc) + iacc)
    r = rng.standard_normal(nc)
    p = ch.poisson(np.asfortranarray(r))
    q = box.poisson(np.asfortranarray(np.concatenate([r, r[:, ::-1, :]], axis=1)))
    assert np.abs(q[:, :ny, :] - q[:, ::-1, :][:, :ny, :]).max() < 1e-12 * np.abs(q).max()
    assert np.abs(p - q[:, :ny, :]).max() < 1e-12 * np.abs(q).max(), np.abs(p - q[:, :ny, :]).max()


def test_tgv_early_viscous_decay(po):
    """Taylor-Green vortex (src/initialisation.f90:892-975): for t << 1 the
    kinetic energy decays at the viscous rate 2 nu |k|^2 = 6/Re."""
    from chapsim2_amd.initial import tgv_fields

    nc = (32, 32, 32)
    L = 2 * np.pi
    f = po.Flow(list(nc), L, -np.pi, np.pi, L, istret=0, wall_y=False, iacc=4, poisson=3, dt=1e-3,
                ren=1600.0, idriven=0)
    f.set_state(*tgv_fields(nc, L, -np.pi, np.pi, L))
    f.update_re(0)
    f.substep(0)

    def ke():
        qx, qy, qz, _ = f.get_state()
        return 0.5 * np.mean(qx**2 + qy**2 + qz**2)

    e0 = ke()
    for it in range(1, 11):
        f.step(it)
    rate = np.log(ke() / e0) / 0.01
    assert abs(rate + 6.0 / 1600.0) < 1e-4 * 6.0 / 1600.0
    assert np.abs(f.divergence()).max() < 1e-13
