"""The oracle's CFL diagnostics (Check_cfl_convection / Check_cfl_diffusion,
src/tools_solver.f90:220-375) against closed forms."""
import numpy as np

LX, LZ = 4 * np.pi, 4 * np.pi / 3


def test_cfl_of_uniform_flow(po):
    """A uniform (U, 0, W) flow is interpolated exactly, so the convective
    CFL number is |U dt/dx + W dt/dz| (the x and z contributions add with
    their signs); on a uniform y grid the diffusion number is
    (1/dx^2 + 1/dz^2 + 1/dy^2) 2 dt / Re."""
    nc = (16, 12, 8)
    dt, re = 2e-3, 400.0
    f = po.Flow(list(nc), LX, -1.0, 1.0, LZ, istret=0, wall_y=False, iacc=4, poisson=2, dt=dt, ren=re)
    for U, W in ((1.5, 0.25), (-0.75, 2.0)):
        qx = np.full(f.shapes["qx"], U, order="F")
        qy = np.zeros(f.shapes["qy"], order="F")
        qz = np.full(f.shapes["qz"], W, order="F")
        pr = np.zeros(f.shapes["pr"], order="F")
        f.set_state(qx, qy, qz, pr)
        f.update_re(1)
        conv, diff = f.cfl()
        dx, dy, dz = LX / nc[0], 2.0 / nc[1], LZ / nc[2]
        assert abs(conv - abs(U * dt / dx + W * dt / dz)) < 1e-13
        assert abs(diff - (1 / dx**2 + 1 / dz**2 + 1 / dy**2) * 2 * dt / re) < 1e-13 * diff


def test_cfl_of_wall_normal_flow_on_stretched_grid(po):
    """A uniform wall-normal velocity V on the stretched channel grid: the
    cell-centred value is V away from the walls, and the CFL number is the
    largest V dt / dy_c, dy_c = h2 / yMappingcc (yMappingcc = ds/dy at the
    centres), attained at the finest cells next to the walls."""
    nc = (16, 24, 8)
    dt, re = 5e-4, 2800.0
    f = po.Flow(list(nc), LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, wall_y=True, iacc=4, poisson=2, dt=dt,
                ren=re, uwall=((0.0, 0.7, 0.0), (0.0, 0.7, 0.0)))
    V = 0.7
    qx = np.zeros(f.shapes["qx"], order="F")
    qy = np.full(f.shapes["qy"], V, order="F")
    qz = np.zeros(f.shapes["qz"], order="F")
    f.set_state(qx, qy, qz, np.zeros(f.shapes["pr"], order="F"))
    f.update_re(1)
    conv, diff = f.cfl()
    h2 = 1.0 / nc[1]
    ymc = f.dom.ymc
    assert abs(conv - V * dt * ymc.max() / h2) < 1e-12 * conv
    yp = f.dom.yp
    assert abs(diff - ((nc[0] / LX) ** 2 + (nc[2] / LZ) ** 2 + (1 / np.diff(yp).min()) ** 2) * 2 * dt / re) < 1e-12 * diff


def test_bulk_monitor_of_uniform_flow(po):
    """For a uniform (U, V, W) flow (wall values V) the interpolated
    velocity is exact, so the bulk mean kinetic energy is (U^2 + V^2 + W^2)/2
    and the bulk qx is U, on the stretched channel grid (the volume weights
    h2 / yMappingcc sum to the channel height)."""
    nc = (16, 24, 8)
    U, V, W = 1.25, -0.5, 0.75
    f = po.Flow(list(nc), LX, -1.0, 1.0, LZ, istret=2, rstret=0.075, wall_y=True, iacc=4, poisson=2, dt=5e-4,
                ren=2800.0, uwall=((0.0, V, 0.0), (0.0, V, 0.0)))
    f.set_state(np.full(f.shapes["qx"], U, order="F"), np.full(f.shapes["qy"], V, order="F"),
                np.full(f.shapes["qz"], W, order="F"), np.zeros(f.shapes["pr"], order="F"))
    mke, qx = f.bulk()
    assert abs(mke - 0.5 * (U * U + V * V + W * W)) < 1e-13
    assert abs(qx - U) < 1e-13


def _window_reference(nc):
    """max |div| of qx = cos(2 pi x / LX) (qy = qz = 0) over the three x
    windows of Check_element_mass_conservation (1:4, 5:n-4, n-5:n,
    src/eq_continuity.f90:387-389): the CP6 staggered first derivative of a
    Fourier mode is -T(kh)/h sin(k x_c) with the scheme's transfer function
    T(kh) = (2a sin(kh/2) + 2b sin(3kh/2)) / (1 + 2 alpha cos kh), alpha =
    9/62, a = 63/62, b = 17/186 (the textbook staggered sixth-order scheme)."""
    n = nc[0]
    h = LX / n
    k = 2 * np.pi / LX
    xf = np.arange(n) * h
    al, a, b = 9 / 62, 63 / 62, 17 / 186
    kh = k * h
    T = (2 * a * np.sin(kh / 2) + 2 * b * np.sin(1.5 * kh)) / (1 + 2 * al * np.cos(kh))
    d = np.abs(T / h * np.sin(k * (xf + h / 2)))
    qx = np.asfortranarray(np.broadcast_to(np.cos(k * xf)[:, None, None], (n, nc[1], nc[2])))
    return qx, [d[0:4].max(), d[4:n - 4].max(), d[n - 6:n].max()], d


def test_mass_check_windows_closed_form(po):
    """The oracle's Check_element_mass_conservation on qx = cos(2 pi x / LX)
    alone: each window's max |div| is the closed form over exactly its own x
    range.  On 48 cells the windows' maxima sit on their inner ends, so a
    window one cell longer or shorter changes the result (checked on the
    closed form itself)."""
    nc = (48, 8, 8)
    f = po.Flow(list(nc), LX, -1.0, 1.0, LZ, istret=0, wall_y=True, iacc=4, poisson=2, dt=1e-3, ren=100.0)
    qx, ref, d = _window_reference(nc)
    n = nc[0]
    assert d[0:5].max() > ref[0] * 1.1 and d[0:3].max() < ref[0] / 1.1
    assert d[n - 7:n].max() > ref[2] * 1.05 and d[n - 5:n].max() < ref[2] / 1.05
    f.set_state(qx, np.zeros(f.shapes["qy"], order="F"), np.zeros(f.shapes["qz"], order="F"),
                np.zeros(f.shapes["pr"], order="F"))
    out = f.mass_check()
    assert np.allclose(out, ref, rtol=1e-12, atol=0), (out, ref)
