"""Initial fields of the benchmark cases.

* Channel: laminar Poiseuille profile 1.5 (1 - y^2/a^2) on cell centres plus
  uniform noise of the given amplitude on every component, wall-normal
  velocity zero on the wall nodes, and qx rescaled to unit bulk velocity
  (src/initialisation.f90:161-247 and :262-414, with the per-point gfortran
  generator replaced by a seeded numpy PCG64 stream; the reference takes such
  fields through its restart files, src/io_restart.f90:112-130).
* Taylor-Green vortex: src/initialisation.f90:892-975.
Arrays are Fortran-ordered float64 with the staggered extents of the
reference.
"""
import numpy as np


def tanh_mesh(ny, lyb, lyt, istret, rstret):
    """(yp, yc, ymp, ymc) of src/geometry.f90:151-269 (tanh) or uniform."""
    if istret == 0:
        h = (lyt - lyb) / ny
        yp = np.arange(ny + 1) * h + lyb
        yc = np.arange(ny) * h + h * 0.5 + lyb
        return yp, yc, np.ones(ny + 1), np.ones(ny)
    gamma, delta, ymin, ymax = {2: (0.5, 0.5, -1.0, 1.0), 3: (-1.0, 1.0, 0.0, 1.0),
                                4: (1.0, 0.0, 0.0, 1.0)}[istret]
    beta = rstret * 20.0
    mm = np.tanh(beta * gamma)
    ff = (lyt - lyb) / (ymax - ymin)

    def one(n, node):
        shift = 0.0 if node else 1.0 / n * 0.5
        step = 1.0 / (n - 1) if node else 1.0 / n
        eta = shift + np.arange(n) * step
        s = np.tanh(beta * (eta - delta)) / mm
        q = 1.0 - s * s * mm * mm
        q = np.where(np.abs(q) < 1e-20, 1.0, mm / beta / np.where(q == 0, 1.0, q))
        return (s - ymin) * ff + lyb, q / ff

    yp, ymp = one(ny + 1, True)
    yc, ymc = one(ny, False)
    return yp, yc, ymp, ymc


def volume_average_x(var, h, ymc, stretched):
    """Get_volumetric_average_3d_for_var_xcx (src/tools_general.f90:1202-1268)
    of an x-pencil cell array var(n1, ny, nz)."""
    n1, ny, nz = var.shape
    dx, dy0, dz = h
    dy = dy0 / ymc if stretched else np.full(ny, dy0)
    w = (dy[None, :, None] * dx) * dz
    vol = float(np.sum(np.broadcast_to(w, var.shape)))
    return float(np.sum(var * dy[None, :, None] * dx * dz)) / vol


def channel_fields(nc, lxx, lyb, lyt, lzz, istret, rstret, noise=0.1, seed=20261202):
    nx, ny, nz = nc
    yp, yc, ymp, ymc = tanh_mesh(ny, lyb, lyt, istret, rstret)
    rng = np.random.Generator(np.random.PCG64(seed))
    qx = noise * rng.uniform(-1.0, 1.0, size=(nz, ny, nx)).transpose(2, 1, 0)
    qy = noise * rng.uniform(-1.0, 1.0, size=(nz, ny + 1, nx)).transpose(2, 1, 0)
    qz = noise * rng.uniform(-1.0, 1.0, size=(nz, ny, nx)).transpose(2, 1, 0)
    qy[:, 0, :] = 0.0
    qy[:, -1, :] = 0.0
    a = (yp[-1] - yp[0]) * 0.5
    prof = (1.0 - (yc ** 2) / a / a) * 1.5
    qx = qx + prof[None, :, None]
    h = (lxx / nx, 1.0 / ny if istret else (lyt - lyb) / ny, lzz / nz)
    ub = volume_average_x(qx, h, ymc, istret != 0)
    qx = qx / ub
    pr = np.zeros((nx, ny, nz))
    return [np.asfortranarray(a) for a in (qx, qy, qz, pr)]


def tgv_fields(nc, lxx, lyb, lyt, lzz):
    """3-D Taylor-Green vortex on a uniform periodic mesh."""
    nx, ny, nz = nc
    hx, hy, hz = lxx / nx, (lyt - lyb) / ny, lzz / nz
    xp = np.arange(nx) * hx
    xc = (np.arange(nx) + 0.5) * hx
    yc = np.arange(ny) * hy + hy * 0.5 + lyb
    yp = np.arange(ny) * hy + lyb
    zc = hz * (np.arange(nz) + 0.5)
    X, Y, Z = np.meshgrid(xp, yc, zc, indexing="ij")
    qx = np.sin(X) * np.cos(Y) * np.cos(Z)
    X, Y, Z = np.meshgrid(xc, yp, zc, indexing="ij")
    qy = -np.cos(X) * np.sin(Y) * np.cos(Z)
    qz = np.zeros((nx, ny, nz))
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    pr = 1.0 / 16.0 * (np.cos(2.0 * X) + np.cos(2.0 * Y)) * (np.cos(2.0 * Z) + 2.0)
    return [np.asfortranarray(a) for a in (qx, qy, qz, pr)]


# enthalpy range of the heated channel's property table: rho*h is monotone
# (decreasing) on it and it holds the bulk (0.8) and wall (1.2) enthalpies
HEATED_H_RANGE = (0.65, 1.225)


def heated_channel_table(n=4096, h0=0.0, h1=2.0):
    """Non-dimensional property table of the variable-property heated channel
    case: n rows over h in [h0, h1] with rho = 1/(1 + 0.8 tanh(5(h-1))),
    cp = 1 + 7 exp(-((h-1)/0.05)^2) (a peak of height 8 at h = 1),
    T = 1 + int_0^h dh'/cp (so that dh = cp dT), mu = exp(-0.6 (h-1)) and
    kappa = 1 + 0.25 tanh(3(h-1)) (smooth monotone fits), beta = -rho'(h) cp / rho.
    Rows are (P, H, T, D, M, K, CP, B), the column order of the reference's
    property tables; use it with ref_t0 <= 0 (already non-dimensional).

    rho*h = h rho(h) is not monotonic over [0, 2] (a maximum of 2.640 at
    h = 0.6306, a minimum of 0.7436 at h = 1.2282), so the bisection on rho*h
    of ftp_refresh_thermal_properties_from_DH (src/input_thermo.f90:393-446)
    can land on the wrong branch there (rho*h = 1 gives h = 0.2 instead of
    1.0).  A solver run uses HEATED_H_RANGE, the monotone part that holds
    the heated channel's enthalpies (bulk 0.8, walls 1.2); T is integrated
    from h0."""
    h = np.linspace(h0, h1, n)
    fine = np.linspace(h0, h1, 64 * (n - 1) + 1)
    cpf = 1.0 + 7.0 * np.exp(-((fine - 1.0) / 0.05) ** 2)
    integ = np.concatenate([[0.0], np.cumsum(0.5 * (1.0 / cpf[1:] + 1.0 / cpf[:-1]) * np.diff(fine))])
    T = 1.0 + integ[::64]
    rho = 1.0 / (1.0 + 0.8 * np.tanh(5.0 * (h - 1.0)))
    cp = 1.0 + 7.0 * np.exp(-((h - 1.0) / 0.05) ** 2)
    mu = np.exp(-0.6 * (h - 1.0))
    kap = 1.0 + 0.25 * np.tanh(3.0 * (h - 1.0))
    drho = -rho**2 * 0.8 * 5.0 / np.cosh(5.0 * (h - 1.0)) ** 2
    beta = -drho * cp / rho
    return np.stack([np.ones(n), h, T, rho, mu, kap, cp, beta], axis=1)
