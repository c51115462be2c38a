"""Manufactured fields with closed-form derivatives, and the continuous
momentum right-hand side of the incompressible (constant-property) equations
in the stress form the reference discretises (src/eq_momentum2.f90:798-1507):

  rhs_i = - d_j (u_j u_i) - d_i p + nu d_j (d_j u_i + d_i u_j - 2/3 delta_ij div u)

A field is a sum of separable terms c X(x) Y(y) Z(z) with X, Z a cosine or
sine of a given wavenumber and Y a polynomial, so every mixed derivative is
exact.  Test infrastructure only.
"""
import numpy as np


def _trig(kind, k, order, x):
    """order-th derivative of cos(k x) ('c') or sin(k x) ('s')."""
    # d/dx cycles cos -> -sin -> -cos -> sin; sin -> cos -> -sin -> -cos
    seq = {"c": [(np.cos, 1.0), (np.sin, -1.0), (np.cos, -1.0), (np.sin, 1.0)],
           "s": [(np.sin, 1.0), (np.cos, 1.0), (np.sin, -1.0), (np.cos, -1.0)]}[kind]
    fn, sg = seq[order % 4]
    return sg * (k ** order) * fn(k * x)


class Field:
    def __init__(self, terms):
        """terms: (c, xkind, kx, Ypoly, zkind, kz) with Ypoly a numpy Polynomial."""
        self.terms = list(terms)

    def ev(self, ax, ay, az, x, y, z):
        out = np.zeros((len(x), len(y), len(z)))
        for c, xk, kx, Y, zk, kz in self.terms:
            X = _trig(xk, kx, ax, np.asarray(x, float))
            Yv = Y.deriv(ay)(np.asarray(y, float)) if ay else Y(np.asarray(y, float))
            Z = _trig(zk, kz, az, np.asarray(z, float))
            out += c * X[:, None, None] * Yv[None, :, None] * Z[None, None, :]
        return np.asfortranarray(out)


def channel_rhs(u, v, w, p, nu, at_x, at_y, at_z):
    """Continuous momentum right-hand sides of (u, v, w, p) at the points
    at_x = (x, y, z) of the x component and likewise for y and z."""
    q = (u, v, w)
    out = []
    for i, pts in enumerate((at_x, at_y, at_z)):
        def d(f, *ax):
            a = [0, 0, 0]
            for s in ax:
                a[s] += 1
            return f.ev(a[0], a[1], a[2], *pts)

        qi = d(q[i])
        conv = np.zeros_like(qi)
        for j in range(3):
            # d_j (u_j u_i) = (d_j u_j) u_i + u_j d_j u_i
            conv += d(q[j], j) * qi + d(q[j]) * d(q[i], j)
        div_i = sum(d(q[j], j, i) for j in range(3))  # d_i (div u)
        visc = np.zeros_like(qi)
        for j in range(3):
            visc += d(q[i], j, j) + d(q[j], i, j)
        visc -= 2.0 / 3.0 * div_i
        out.append(-conv - d(p, i) + nu * visc)
    return out


def channel_rhs_var(u, v, w, p, nu, rho, mu, at_x, at_y, at_z):
    """Continuous momentum right-hand sides with variable density and
    viscosity, the form of the reference's thermo path:

      rhs_i = - d_j (rho u_j u_i) - d_i p + nu d_j [mu (d_j u_i + d_i u_j - 2/3 delta_ij div u)]

    rho and mu are functions of y alone, given as (f, f') pairs of callables."""
    q = (u, v, w)
    out = []
    for i, pts in enumerate((at_x, at_y, at_z)):
        def d(f, *ax):
            a = [0, 0, 0]
            for s in ax:
                a[s] += 1
            return f.ev(a[0], a[1], a[2], *pts)

        y = np.asarray(pts[1], float)[None, :, None]
        r, r1 = rho[0](y), rho[1](y)
        m, m1 = mu[0](y), mu[1](y)
        qi = d(q[i])
        conv = np.zeros_like(qi)
        for j in range(3):
            conv += r * (d(q[j], j) * qi + d(q[j]) * d(q[i], j))
        conv += r1 * d(q[1]) * qi  # (d_y rho) u_y u_i
        div = sum(d(q[j], j) for j in range(3))
        div_i = sum(d(q[j], j, i) for j in range(3))
        visc = np.zeros_like(qi)
        for j in range(3):
            visc += d(q[i], j, j) + d(q[j], i, j)
        visc -= 2.0 / 3.0 * div_i
        s_iy = d(q[i], 1) + d(q[1], i) - (2.0 / 3.0 * div if i == 1 else 0.0)
        out.append(-conv - d(p, i) + nu * (m * visc + m1 * s_iy))
    return out


class EtaProfile:
    """y profile g(eta(y)) on the two-sided tanh mesh (src/geometry.f90:
    151-269, istret = 2, y in [-1, 1]): eta = 1/2 + atanh(m y) / beta with
    m = tanh(beta / 2) is the uniform computational coordinate, g a
    polynomial in s = 2 eta - 1.  deriv(k) gives d^k/dy^k (k <= 2) by the
    chain rule."""

    def __init__(self, g, beta):
        self.g, self.beta = g, float(beta)
        self.m = np.tanh(self.beta * 0.5)

    def _eta(self, y):
        m, b = self.m, self.beta
        e = 0.5 + np.arctanh(m * y) / b
        e1 = m / (b * (1.0 - (m * y) ** 2))
        e2 = 2.0 * m ** 3 * y / (b * (1.0 - (m * y) ** 2) ** 2)
        return e, e1, e2

    def __call__(self, y):
        e, _, _ = self._eta(np.asarray(y, float))
        return self.g(2.0 * e - 1.0)

    def deriv(self, k):
        if k == 0:
            return self
        g1, g2 = self.g.deriv(1), self.g.deriv(2)

        def f(y):
            e, e1, e2 = self._eta(np.asarray(y, float))
            s = 2.0 * e - 1.0
            if k == 1:
                return 2.0 * g1(s) * e1
            if k == 2:
                return 4.0 * g2(s) * e1 * e1 + 2.0 * g1(s) * e2
            raise ValueError("k <= 2")

        return f
